/* xted: MI355X-native parallel Zhang-Shasha tree edit distance, C ABI.
 *
 * This is the thin C layer the C++ drop-in (include/TED_C++.h) and any other
 * host language bind to. Plain pointers and sizes only; no exceptions cross
 * it; every entry point returns an xted status code and the thread's last
 * error message is available from xted_last_error().
 *
 * Tree encoding: n nodes in preorder, parent[0] = -1, 0 <= parent[i] < i and
 * parent[i] is an ancestor-or-self of node i-1; label[i] is an int32 label
 * id (equal ids are equal labels, unit costs). The reference passes the same
 * trees as preorder children lists plus string labels
 * (X-TED_GPU/TED_C++.h:27-28, README.md:31-44); the C++ drop-in converts.
 *
 * Tables: td is m x n row-major in preorder, td[i*n + j] =
 * TED(subtree_x(i), subtree_y(j)); the distance is td[0].
 */
#ifndef XTED_H
#define XTED_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define XTED_OK 0
#define XTED_ERR_INVALID_ARG 1
#define XTED_ERR_NOT_PREORDER 2
#define XTED_ERR_OUT_OF_MEMORY 3
#define XTED_ERR_DEVICE 4
#define XTED_ERR_TOO_LARGE 5
#define XTED_ERR_PARSE 6
#define XTED_ERR_BUSY 7
#define XTED_ERR_TIMEOUT 8
#define XTED_ERR_NOT_FOUND 9
#define XTED_ERR_CANCELLED 10

/* Granularity classes: 0 = tables with a leaf keyroot (closed form, one
 * launch per batch); 1 + wi*5 + ci = workgroup row scan with 2^wi waves and
 * 2^ci columns per lane (wi, ci in 0..4; the classes xted_force_class
 * takes); 26 = no form fits (XTED_ERR_TOO_LARGE); 27..32 = column blocks
 * of 4096, 2048, 1024, 512, 256 or 128 columns, one workgroup per block, for
 * tables too wide or too LDS-hungry for one workgroup (the role of the
 * reference's multi-block and last-table kernels,
 * X-TED_GPU/parallel_X_TED.cu:883-1131); 33, 34, 35 = tables of at most 16,
 * 32 or 64 columns, four, two or one to a wave (the role of the reference's
 * single-thread tasks, X-TED_GPU/parallel_X_TED.cu:382). */
#define XTED_CLASS_AUTO (-1)
#define XTED_CLASS_LANE 0

typedef struct xted_context xted_context;

typedef struct xted_stats {
  double plan_ms;   /* device plan build incl. the bucket-count read-back */
  double dp_ms;     /* forest-DP level loop (device time) */
  double total_ms;  /* whole call, host inputs to host outputs */
  int64_t cells;    /* interior forest cells, sum over pairs of S_x * S_y */
  int64_t tables;   /* keyroot-pair tables */
  int32_t levels;
  int32_t launches; /* forest kernel launches */
  /* column-blocked tables: cross-workgroup reads issued ahead that were
   * stale and had to be polled again, and pipeline checks that waited */
  int64_t xl_stale_reads;
  int64_t xl_waits;
} xted_stats;

/* ---- context --------------------------------------------------------------
 * One context = one GPU, one HIP stream, device buffers reused across calls.
 * Calls on one context are serialised; use one context per host thread for
 * concurrency. The reference keeps equivalent state in globals
 * (X-TED_GPU/parallel_X_TED.cu:59-65, X-TED_GPU/TED.cu:3). */
int xted_create(int32_t device, xted_context** out);
int xted_destroy(xted_context* ctx);
const char* xted_last_error(void);
const char* xted_status_string(int32_t status);

/* ---- drop-in single pair ----------------------------------------------------
 * Replaces standard_ted_1 / standard_ted (X-TED_GPU/TED.cu:78-128) and the
 * scheduler behind them, parallel_standard_ted_test
 * (X-TED_GPU/parallel_X_TED.cu:169-399). td (m*n ints) may be NULL.
 * Empty trees follow the reference: m == 0 gives n, n == 0 gives m
 * (X-TED_GPU/TED.cu:119-124). The pair goes through the context's tree
 * pool, which this call (and xted_edit_mapping) clears first: keep pools
 * that batches or a server use on contexts of their own. */
int xted_ted(xted_context* ctx, const int32_t* x_parent, const int32_t* x_label, int32_t m,
             const int32_t* y_parent, const int32_t* y_label, int32_t n, int32_t* td,
             int32_t* dist, xted_stats* stats);

/* Optimal edit mapping of x (m nodes) and y (n nodes): the tree-distance
 * table is computed on the device as by xted_ted, then an optimal mapping is
 * traced back on the host through the forest tables of the tree pair and of
 * the subtree pairs it matches as units. map_x[k], map_y[k] (capacity
 * min(m, n) each) receive the *nmap matched node pairs in increasing order of
 * both; unmatched x nodes are deleted, unmatched y nodes inserted, matched
 * pairs with different labels renamed, and the cost of that script is
 * *dist. No counterpart in the reference, which returns distances only;
 * m * n at most 2^26 (XTED_ERR_TOO_LARGE beyond). Unit costs only: a
 * context under a cost model (xted_set_costs) returns XTED_ERR_INVALID_ARG. */
int xted_edit_mapping(xted_context* ctx, const int32_t* x_parent, const int32_t* x_label,
                      int32_t m, const int32_t* y_parent, const int32_t* y_label, int32_t n,
                      int32_t* map_x, int32_t* map_y, int32_t* nmap, int32_t* dist);
/* The host half of xted_edit_mapping, for a tree-distance table the caller
 * already holds (xted_ted or xted_batch output): td is m x n, row-major.
 * *cost receives the mapping's cost (td[0] for a correct table). Runs on the
 * host only. */
int xted_mapping_from_td(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                         const int32_t* y_parent, const int32_t* y_label, int32_t n,
                         const int32_t* td, int32_t* map_x, int32_t* map_y, int32_t* nmap,
                         int64_t* cost);

/* ---- structural arrays --------------------------------------------------------
 * Device-computed outermost-right leaves (orl, n ints), keyroots in
 * descending preorder (kr, up to n ints, count in *K) and keyroot depth per
 * keyroot (kr_depth, K ints). Replaces outermost_right_leaves
 * (X-TED_GPU/TED.cu:21-42), key_roots (X-TED_GPU/TED.cu:44-75) and the
 * keyroot depth loop (X-TED_GPU/parallel_X_TED.cu:184-212). Any output may be
 * NULL. */
int xted_tree_structure(xted_context* ctx, const int32_t* parent, int32_t n, int32_t* orl,
                        int32_t* kr, int32_t* K, int32_t* kr_depth);

/* ---- device-resident tree pool and batches ------------------------------------
 * No counterpart in the reference, which evaluates one pair per call
 * (X-TED_GPU/test_TED.cu:36-51). Trees are concatenated: tree t has sizes[t]
 * nodes. ids receives the pool ids (may be NULL). */
int xted_pool_add(xted_context* ctx, int32_t ntrees, const int32_t* sizes,
                  const int32_t* parents, const int32_t* labels, int32_t* ids);
int xted_pool_clear(xted_context* ctx);
int xted_pool_size(xted_context* ctx, int32_t* ntrees);
/* Distances of the pairs (x_ids[i], y_ids[i]) of pool trees, all pairs in one
 * device plan. td_out (may be NULL; entries may be NULL) receives each pair's
 * m x n table. */
int xted_batch(xted_context* ctx, int32_t npairs, const int32_t* x_ids, const int32_t* y_ids,
               int32_t* dist, int32_t* const* td_out, xted_stats* stats);
/* xted_batch without device allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Every
 * device buffer the batch needs (plan arrays, TD arena, distances) is carved
 * from one fixed array of XTED_NOALLOC_BYTES (64 MiB) that the library holds
 * on each device it runs on until the process ends; the call makes no
 * hipMalloc / hipFree / pinned-host allocation and instantiates no graph
 * (its levels are launched one by one). Pairs whose buffers exceed the array
 * (8*m*n bytes of tables plus 72 bytes per planned keyroot-pair table, as in
 * xted_all_pairs) are refused with XTED_ERR_TOO_LARGE. Distances only. Calls
 * on one context are serialised; calls on different contexts of one device
 * share that device's array and must not overlap. */
#define XTED_NOALLOC_BYTES (64 << 20)
int xted_batch_noalloc(xted_context* ctx, int32_t npairs, const int32_t* x_ids,
                       const int32_t* y_ids, int32_t* dist);
/* All-pairs / dataset join: the ntrees x ntrees distance matrix of a set of
 * trees, dist[i*ntrees + j] = TED(tree i, tree j) (trees concatenated as in
 * xted_pool_add). The reference's test() sizes a dimension x dimension
 * Distance matrix but fills only (0, 1) (X-TED_GPU/test_TED.cu:30-54). The
 * pairs run through the batched plan in chunks whose device bytes take at
 * most max_bytes of HBM (0: 16 GiB): a pair counts its TD and TD^T (8*m*n
 * bytes) and its plan arrays (72 bytes per planned keyroot-pair table: the
 * tables of two non-leaf keyroots with unit costs, all K_x*K_y tables under
 * a cost model). With unit costs the matrix is
 * symmetric with a zero diagonal and only i < j is computed; under a cost
 * model every ordered pair, the diagonal included, is. The context's pool
 * is used as scratch (cleared first). stats sums the chunks. */
int xted_all_pairs(xted_context* ctx, int32_t ntrees, const int32_t* sizes,
                   const int32_t* parents, const int32_t* labels, int64_t max_bytes,
                   int32_t* dist, xted_stats* stats);

/* ---- tuning and test hooks ------------------------------------------------------
 * Granularity cut-overs, the role of the fixed thresholds 15 / 18,750 /
 * 25,250,000 in X-TED_GPU/parallel_X_TED.cu:382. lds_per_wave: slot bytes
 * one wave may own before a table gets more waves; max_cpl (1, 2, 4, 8 or
 * 16): columns per lane before a table gets more waves (at most 8 under a
 * cost model, whose 16-column forms run at 2 waves per SIMD); tables with at
 * least wide_cols columns get at least min_waves waves. */
int xted_set_tuning(xted_context* ctx, int32_t lds_per_wave, int32_t max_cpl,
                    int32_t min_waves, int32_t wide_cols);
/* Forces every table that fits into class cls, a row-scan class (1..25), a
 * column-blocked one (27..32) or a tiny form (33..35); XTED_CLASS_AUTO
 * restores the rule.
 * lane_enable = 0 sends single-row tables to the row-scan kernels. */
int xted_force_class(xted_context* ctx, int32_t cls, int32_t lane_enable);
/* Levels of a batch that at most tail_pairs pairs reach (the last levels of
 * the deepest pairs) hold few tables, so their time is one table's latency:
 * there the rule takes the fewest waves, up to 16 columns per lane and slot
 * storage up to the device limit. 0 turns this off. */
int xted_set_tail_rule(xted_context* ctx, int32_t tail_pairs);
/* Column-blocked tables with at most seq_rows rows are swept block after
 * block by one workgroup (no waits between workgroups); taller ones run one
 * workgroup per block, chained per row. */
int xted_set_xl_rule(xted_context* ctx, int32_t seq_rows);
/* Widest column block (128 to 4096 columns; default 4096) that a taller
 * column-blocked table may use: narrower blocks put more workgroups on one
 * table. The reference's multi-block form uses 8 blocks of 1024 threads
 * (X-TED_GPU/parallel_X_TED.cu:883-961). Results do not depend on it. */
int xted_set_xl_width(xted_context* ctx, int32_t max_block_cols);
/* Streams (1 to 4) that the launches of one level run on concurrently. The
 * reference gives each class kernel its own stream but waits for each
 * launch before the next (X-TED_GPU/parallel_X_TED.cu:433-438). Results do
 * not depend on it. */
int xted_set_level_streams(xted_context* ctx, int32_t n);
/* A batch of at least split_min pairs is cut into groups (1 to 4; pair i in
 * group i mod groups), each planned and run on its own stream: one group's
 * level barriers and last tables overlap the others' work, and a group's
 * plan overlaps the levels of the groups before it. No counterpart in the
 * reference, which runs one pair at a time. Results do not depend on it. */
int xted_set_batch_groups(xted_context* ctx, int32_t groups, int32_t split_min);
/* Where each pair's levels sit among the batch's levels: 0 (the default)
 * starts every pair at level 0, as the reference's single pair does
 * (X-TED_GPU/parallel_X_TED.cu:214-220); 1 ends every pair at the batch's
 * last level; 2 places it midway. A pair's own levels keep their order
 * (level = depth_x + depth_y + offset), so results do not depend on it; the
 * offsets move the tables of shallow pairs into the levels whose time is set
 * by the few large tables of the deepest pairs. */
int xted_set_level_align(xted_context* ctx, int32_t mode);
/* mode 1 (the default): the tables of each launch start in four size
 * bands, longest first (by rows), so a launch ends on short tables; 2: in
 * addition, the row-scan tables of a level and class that need more slot
 * storage than a workgroup's share of LDS at full occupancy take a launch of
 * their own, so they do not lower the occupancy of the others; 0: one
 * launch per level and class, tables in plan order. The reference sorts
 * each level's queue by table size, ascending (thrust::sort_by_key,
 * X-TED_GPU/parallel_X_TED.cu:374). Results do not depend on it. Other
 * modes: XTED_ERR_INVALID_ARG. */
int xted_set_size_bands(xted_context* ctx, int32_t mode);
/* Size-band mode 2: a row-scan table of a W-wave form is heavy when its slot
 * rows need more than W * bytes_per_wave bytes of LDS (default 10240: a
 * workgroup's share at 16 waves per CU). Bytes outside [1024, 153600]:
 * XTED_ERR_INVALID_ARG. Results do not depend on it. */
int xted_set_light_lds(xted_context* ctx, int32_t bytes_per_wave);
/* Tiny forms in use: bit k (k = 0, 1, 2) lets tables of at most 16 << k
 * columns whose slot rows fit in 2048 << k bytes take class 33 + k, one table
 * per (16 << k)-lane segment of a wave; a table takes the narrowest form in
 * use that holds it, and other tables take the row-scan forms.
 * Default 1 (16 columns). The reference runs every table with one
 * thread per cell anti-diagonal (X-TED_GPU/parallel_X_TED.cu:883-1131).
 * Results do not depend on it. A mask outside [0, 7]: XTED_ERR_INVALID_ARG. */
int xted_set_tiny_forms(xted_context* ctx, int32_t mask);
/* on = 1 (the default): the level launches of a batch without
 * column-blocked tables are captured once as a HIP graph and re-launched
 * while a later batch plans to the same launches (same counts per level and
 * class, same buffers); 0 launches them one by one. Profiling launches them
 * one by one regardless. The reference launches every kernel directly
 * (X-TED_GPU/parallel_X_TED.cu:429-443). Results do not depend on it. */
int xted_set_graphs(xted_context* ctx, int32_t on);
/* Non-unit edit costs (the reference hard-codes unit costs,
 * X-TED_CPU/TED.cpp:111,115,122): labels in [0, alphabet) (alphabet at most
 * 4096), del[a] / ins[a] for deleting / inserting a node labelled a and
 * ren[a*alphabet + b] for relabelling a (an x label) to b (a y label), each
 * in [0, 32767]. Applies to every later xted_ted, xted_batch and server
 * query of the context; the tables of a leaf keyroot then take the row-scan
 * and tiny forms (their closed form assumes unit costs), and a pair
 * whose forest values can exceed 65535 (deleting all of x plus inserting
 * all of y) is rejected with XTED_ERR_TOO_LARGE. xted_ted of an empty tree
 * returns the cost of inserting (deleting) every node of the other.
 * del == NULL restores unit costs. The call first waits for the work the
 * context has launched; server queries planned under the previous model and
 * not yet fully launched then fail with XTED_ERR_INVALID_ARG. Under a cost
 * model xted_pool_add rejects
 * a tree with a label outside the alphabet; a tree added before the model
 * stays in the pool, and a query that uses one fails with
 * XTED_ERR_INVALID_ARG. */
int xted_set_costs(xted_context* ctx, int32_t alphabet, const int32_t* del, const int32_t* ins,
                   const int32_t* ren);

/* ---- per-class device timing ---------------------------------------------------
 * With profiling on, xted_ted / xted_batch bracket every forest-DP launch
 * with events on the context's stream; xted_class_profile reports, for class
 * cls of the last call, the launches, tables, interior cells and summed
 * device milliseconds. The reference times its kernels the same way
 * (X-TED_GPU/parallel_X_TED.cu:429-443). */
int xted_set_profile(xted_context* ctx, int32_t on);
int xted_class_profile(xted_context* ctx, int32_t cls, int64_t* launches, int64_t* tables,
                       int64_t* cells, double* ms);
/* Launch i of the last profiled call, in launch order: its level (-1 for the
 * leaf kernel, which precedes level 0), class, tables, interior cells, start
 * (ms after the call's first timed launch started) and device duration.
 * *count receives the number of timed launches; i < 0 only sets it. With the
 * level loop of X-TED_GPU/parallel_X_TED.cu:312-387 this gives each level's
 * span and its longest launch (the batch's critical path). */
int xted_launch_profile(xted_context* ctx, int32_t i, int32_t* level, int32_t* cls,
                        int64_t* tables, int64_t* cells, double* start_ms, double* ms,
                        int32_t* count);

/* ---- continuous-batching server ---------------------------------------------------
 * No counterpart in the reference. A server thread drains submitted queries
 * over pool trees into device batches, admitting each by the HBM its TD and
 * TD^T tables (8*m*n bytes) and its plan arrays (72 bytes per planned
 * keyroot-pair table, as in xted_all_pairs) need against hbm_budget bytes. */
typedef struct xted_server xted_server;
typedef struct xted_result {
  uint64_t tag;         /* caller's tag from submit */
  int32_t status;       /* XTED_OK or an error code */
  int32_t dist;
  double latency_ms;    /* submit -> result available */
  double queue_ms;      /* submit -> admitted into a device batch */
} xted_result;
int xted_server_create(xted_context* ctx, int64_t hbm_budget, int32_t max_batch,
                       xted_server** out);
int xted_server_submit(xted_server* srv, int32_t x_id, int32_t y_id, uint64_t tag);
/* Copies up to cap finished results, waiting up to timeout_ms for at least
 * one (0 = do not wait). *count receives the number copied. */
int xted_server_poll(xted_server* srv, xted_result* out, int32_t cap, int32_t timeout_ms,
                     int32_t* count);
int xted_server_destroy(xted_server* srv);

/* ---- host utilities (no GPU needed) ---------------------------------------------------
 * Seeded synthetic trees: kind 0 = random recursive tree, 1 = deep/narrow,
 * 2 = shallow/wide; depth_min/depth_max > 0 reject trees outside the band. */
/* Text formats (host only). Bracket notation, one tree per line as
 * {label{child}{child}...}: the reference ships such files
 * (Sampled_Dataset_Extra_Large_Trees/synthetic/<n>_nodes/synthetic_tree_<n>.txt)
 * but reads none of them. xted_bracket_to_lines converts one tree to the
 * reference's input lines (README.md:31-42): the label line (each label
 * followed by a space) and the preorder adjacency line ("[[1, 2], [], []]").
 * xted_lines_to_bracket converts back. Outputs are NUL-terminated; *_need
 * receives the bytes each needs (with the NUL) and a short buffer gives
 * XTED_ERR_TOO_LARGE with nothing written; malformed text gives
 * XTED_ERR_PARSE. No counterpart in the reference. */
int xted_bracket_to_lines(const char* bracket, char* labels_line, int64_t labels_cap,
                          char* adj_line, int64_t adj_cap, int64_t* labels_need,
                          int64_t* adj_need);
int xted_lines_to_bracket(const char* labels_line, const char* adj_line, char* bracket,
                          int64_t cap, int64_t* need);

int xted_generate(int32_t kind, int32_t n, int32_t alphabet, uint64_t seed, int32_t depth_min,
                  int32_t depth_max, int32_t* parent, int32_t* label);
uint64_t xted_derive_seed(uint64_t config, uint64_t index);
int xted_validate(const int32_t* parent, int32_t n);
/* Host restatements used for planning and reporting: S = sum of keyroot
 * subtree sizes (cells of a pair = S_x * S_y) and the depth. */
int xted_tree_stats(const int32_t* parent, int32_t n, int64_t* ksum, int32_t* depth,
                    int32_t* nkeyroots);

#ifdef __cplusplus
}
#endif
#endif
