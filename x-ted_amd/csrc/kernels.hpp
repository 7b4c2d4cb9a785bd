// Kernel argument blocks and launchers (prep.hip, plan.hip, forest.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace xted {

// Tree preprocessing: one workgroup per tree.
struct PrepArgs {
  int32_t* parent;  // input (already in the pool)
  int32_t* label;
  int32_t* orl;
  int32_t* kdep;
  int32_t* kdn;
  int32_t* lrank;
  int32_t* kr;
  int32_t* krkd;
  int32_t* krsize;
  int32_t* krleaf;
  int32_t* skr;
  int32_t* skd;
  int32_t* ssize;
  int32_t* sleaf;
  int32_t* spre;
  int32_t* kdoff;
  int32_t* rtop;
  NodeRec* crec;
  NodeRec* rrec;
  TreeMeta* trees;
  int32_t* scratch;            // 6*(n+1) ints per tree at scratch_off[b]
  const int64_t* scratch_off;  // per block of this launch
  int32_t first_tree;
};
void launch_tree_prep(const PrepArgs& a, int ntrees, hipStream_t s);

// Cost prefix sums of trees [first_tree, first_tree + ntrees): per node i in
// [0, n], pdel[off + i] / pins[off + i] = the deletion / insertion costs of
// nodes [0, i) by label; totals[2t], totals[2t+1] = the whole tree's.
struct CostPrefixArgs {
  const TreeMeta* trees;
  const int32_t* label;
  const int32_t* del;  // per label
  const int32_t* ins;
  int32_t* pdel;
  int32_t* pins;
  int32_t* totals;     // 2 per tree of this launch
  const int32_t* orl;  // the pool's
  CostRec* rec;        // per node: the sums at i and at orl(i) + 1
  int32_t first_tree;
  int32_t alphabet;    // labels outside [0, alphabet) read the last entry
};
void launch_cost_prefix(const CostPrefixArgs& a, int ntrees, hipStream_t s);

// Plan builder: count, scan and scatter of (level, class) buckets.
struct PlanArgs {
  PoolView pool;
  const PairMeta* pairs;
  const int64_t* blk_start;    // npairs+1: first block of each pair
  // Per pair: added to the level of each of its tables (null: none). The
  // levels of one pair keep their order, so any offset is a valid schedule.
  const int32_t* level_off;
  int32_t npairs;
  int32_t nlevels;
  int32_t lchunk_tables;       // target tables per block
  ClassRule rule;
  // Levels reached by only a few pairs hold few tables, so their time is the
  // latency of one table: tail_level[L] != 0 selects rule_tail there.
  ClassRule rule_tail;
  const uint8_t* tail_level;   // nlevels
  int32_t force_class;         // -1 = automatic; else row-scan tables go to this class if they fit
  int32_t bands;               // 0: every table in band 0 (plan order within a launch)
  int32_t light_per_wave;      // bands 2: heavier tables take launches of their own
  int32_t lane_enable;         // 0 = single-row tables use the row-scan kernels
  const uint8_t* pair_leaf;    // per pair: its lines are in closed form (null: lane_enable)
  unsigned long long* counts;  // per bucket (bucket_of: level, class, band)
  unsigned long long* cells;   // per bucket: interior cells
  unsigned long long* cursor;  // scatter cursors (bucket offsets)
  int32_t* maxlds;             // per bucket: largest slot storage
  uint64_t* tasks;
  TaskDesc* descs;             // one per task, same order
  unsigned int* tickets;       // per bucket: column-blocked launch counters
  int32_t* abort_flag;         // kAbortInts: column-blocked waits gave up, what they waited for,
                               // and their stale-read / wait counts
};
// Blocks a pair occupies in the plan kernels (same formula on both sides).
inline int64_t plan_blocks(int32_t Kx, int32_t Ky, int32_t chunk) {
  const int32_t lper = chunk / (Kx > 0 ? Kx : 1) > 0 ? chunk / (Kx > 0 ? Kx : 1) : 1;
  return (Ky + lper - 1) / lper;
}
void launch_plan(const PlanArgs& a, int64_t nblocks, hipStream_t s);
void launch_gather_dist(const PairMeta* pairs, const td_t* td, int32_t npairs, int32_t* dist,
                        hipStream_t s);
// dst[i] = src[i] for i < n: a table of the arena widened to the int32 a
// caller receives.
void launch_widen_td(const td_t* src, int32_t* dst, int64_t n, hipStream_t s);

// Tables with a leaf keyroot, in closed form. For a single node v and a
// subtree T (unit costs), TED({v}, T) = |T| - [label(v) occurs in T]: map v
// onto a node of T with its label if there is one, insert the rest. So the
// row of TD of a leaf keyroot i of x holds, for every node c of y,
//   TD[i][c] = |sub_y(c)| - [label_x(i) occurs in sub_y(c)],
// and a leaf keyroot j of y likewise gives the column TD[.][j]. In preorder,
// sub(c) = [c, orl(c)], so the test is "the first position >= c holding the
// label is <= orl(c)". These cells are exactly the tree cells of every table
// with a leaf keyroot (the tables of X-TED_CPU/parallel_computing.cpp:4-45
// with a one-node side), and they read nothing, so one launch covers them
// all, independent of every level.
struct LeafArgs {
  PoolView pool;
  const PairMeta* pairs;
  // npairs+1: prefix over pairs of ceil(n_x / 64) + ceil(n_y / 64)
  const int64_t* tile_start;
  int32_t npairs;
  td_t* td;
  // Cost model (costs != 0): per-label delete / insert costs and the rename
  // matrix ren[a * alphabet + b]; the pool's pdel / pins hold their prefix
  // sums over each tree's nodes.
  const int32_t* del;
  const int32_t* ins;
  const int32_t* ren;
  int32_t alphabet;
  int32_t costs;
};
// Under a cost model the lines take the cost form of the closed form: for a
// leaf v of x and a node c of y,
//   TD[v][c] = ins(sub_y c) + min(del(v), min over k in sub_y(c) of
//              ren(v, k) - ins(k)),
// (v deleted, or mapped onto one node k of the subtree; everything else
// inserted), and symmetrically for a leaf of y. Its subtree minima take one
// right-to-left sweep over the other tree with a stack of pending sibling
// values as deep as the tree's kLineStack measure (at most kCostLeafStack;
// pairs with deeper trees plan their leaf tables as row scans).
constexpr int kCostLeafStack = 128;
// One workgroup per tile of 64 consecutive nodes of one side of a pair (its
// leaf keyroots are the lines). Writes cells of other tables' nodes as well
// (see forest.hip), so it must finish before any table of the group runs.
void launch_leaf(const LeafArgs& a, int64_t ntiles, hipStream_t s);

// Ints of a group's column-blocked record (DPArgs::xl_abort).
constexpr int kAbortInts = 16;

// Entries of padding the TD arena keeps before and after every table, so
// the row-scan kernels can issue whole-group loads without bounds branches
// (a group of 8 entries starts at most 7 before column c0 or at c0 itself).
constexpr int kTdPad = 16;

// Forest DP.
struct DPArgs {
  PoolView pool;
  const PairMeta* pairs;
  td_t* td;               // arena holding every TD / TDT
  const uint64_t* tasks;  // this launch's tables
  const TaskDesc* descs;  // and their descriptors
  int64_t ntasks;
  // Column-blocked launches only.
  unsigned long long* xl_words;  // cross-workgroup words of every such task
  const int64_t* xl_off;         // this launch's tasks' offsets into xl_words
  unsigned int* xl_ticket;       // (task, block) counter of this launch, zeroed
  int32_t* xl_abort;             // [0] raised when a wait gives up, [1..6] what it waited for,
                                 // [8] stale prefetched reads, [9] throttle checks that waited
  uint32_t xl_epoch;             // tag of this launch's words (never 0)
  int32_t xl_nbmax;              // workgroups per task in the grid
  int32_t xl_seq_rows;           // tables this short: one workgroup sweeps all blocks
  // Non-unit costs (costs != 0): the pool's pdel / pins prefix sums carry
  // the deletion and insertion costs; ren[a * alphabet + b] relabels a to b.
  const int32_t* ren;
  int32_t alphabet;
  int32_t costs;
};
// Arena layout of the column-blocked tasks of a batch (after the scatter).
struct XLLayoutArgs {
  PoolView pool;
  const PairMeta* pairs;
  const uint64_t* tasks;
  const unsigned long long* counts;
  const unsigned long long* cursor;  // scatter cursors: bucket ends
  int32_t nlevels;
  int64_t* off;                 // per column-blocked task, in bucket order
  unsigned long long* total;    // words
};
void launch_xl_layout(const XLLayoutArgs& a, hipStream_t s);
void launch_xl_zero(unsigned long long* p, int64_t n, hipStream_t s);
// Lets every row-scan kernel request up to lds_limit bytes of dynamic LDS.
hipError_t forest_init_attributes(int lds_limit);
// Launches the row-scan kernel of class `cls` over a.tasks[0..ntasks).
// Returns false when the class has no kernel (the leaf class runs as one
// launch_leaf, class kClassXL has none). Column-blocked classes launch
// a.ntasks * a.xl_nbmax workgroups.
bool launch_forest(int cls, const DPArgs& a, int lds_bytes, hipStream_t s);
// Tiny tables of form k (class kClassTiny + k), 256 / tiny_seg(k) per
// workgroup; slot_bytes per table.
void launch_tiny(int form, const DPArgs& a, int slot_bytes, hipStream_t s);

}  // namespace xted
