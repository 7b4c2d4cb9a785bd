// Device-side data layout shared by the HIP kernels and the host engine.
//
// Tree pool (structure of arrays, all int32, one region of n+1 entries per
// tree starting at node_off):
//   node arrays    parent/label/orl/kdep/kdn  [node_off + i], i in [0, n)
//   leaf rank      lrank                      [node_off + i], i in [0, n]
//   keyroot arrays kr/krkd/krsize/krleaf      [node_off + k], k in [0, K)
//                  (keyroots in descending preorder; K <= n)
//   sorted keyroots skr/skd/ssize/sleaf       [node_off + k], k in [0, K):
//                  the K1 leaf keyroots first, then the others by ascending
//                  nesting height d (and size band within a height), those
//                  of height d at [K1 + kdoff[d], K1 + kdoff[d+1])
//   spre           [node_off + k], k in [0, K]: prefix sums of ssize
//   kdoff          [node_off + d], d in [0, kdmax+1]
// A pair owns a tree-distance table TD (m x n, row-major, preorder) and its
// transpose TDT (n x m); both live in one arena of 16-bit entries (td_t) at
// td_off / tdt_off.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define XTED_HD __host__ __device__
#else
#define XTED_HD
#endif

namespace xted {

struct TreeMeta {
  int64_t node_off;  // offset of the tree's region (n+1 entries)
  int32_t n;         // node count
  int32_t K;         // keyroot count
  int32_t kdmax;     // keyroot nesting height of the root
  int32_t nleaves;   // leaf count
  int32_t K1;        // keyroots that are leaves
  int32_t pad;
};

// One entry of a tree-distance table in the arena. Every entry is at most
// the cost of deleting all of x and inserting all of y: with unit costs
// n_x + n_y <= 65534 (pool trees have at most 32767 nodes), under a cost
// model at most 65535 (checked per pair on the host). So 16 bits hold every
// value exactly, and each forest cell reads 2 bytes of HBM instead of 4;
// the kernels compute in int32 and the tables handed to a caller are int32.
using td_t = uint16_t;

struct PairMeta {
  int32_t tx, ty;   // tree ids (rows = x, columns = y)
  int64_t td_off;   // TD  offset (m*n entries)
  int64_t tdt_off;  // TDT offset (n*m entries)
};

// What a forest kernel needs to start one table, written by the plan next to
// the table's task word so that the kernel starts from one 64-byte read
// instead of a chain of dependent reads (task, pair, trees, subtree ends).
// The table is oriented with its short side as rows.
struct alignas(16) TaskDesc {
  int64_t tdo;       // offset of TD in the row layout ([r][c])
  int64_t tdt;       // and in the column layout ([c][r])
  int64_t offR;      // node offset of the row tree in the pool
  int64_t offC;      // and of the column tree
  int32_t ldo, ldt;  // row lengths of the two layouts
  int32_t r0, c0;    // the table's keyroots (row tree, column tree)
  int32_t rmax, cmax;
  int32_t base_rx;   // kdep of r0; bit 31 set when the rows are x's nodes
  int32_t lbase;     // leaves of the column tree before c0
};
static_assert(sizeof(TaskDesc) == 64, "TaskDesc is one 64-byte read");

// What a forest kernel reads of one node, in one 16-byte load: as a column
// of a table (crec: x = leaf rank at orl(i), y = leaf rank of i-1) or as a
// row (rrec: x = kdep(i), y = kdep(i-1)); bit 31 of y is set when i-1 is a
// leaf. Entry 0's y is 0.
struct alignas(16) NodeRec {
  int32_t orl, label, x, y;
};

// Under a cost model, a node's deletion / insertion cost prefix sums at
// itself and just past its subtree (at orl(i) + 1), in one 16-byte load.
struct alignas(16) CostRec {
  int32_t del, ins, del_end, ins_end;
};

// Pointers to the pool arrays (device addresses).
struct PoolView {
  const int32_t* parent;
  const int32_t* label;
  const int32_t* orl;
  const int32_t* kdep;   // nesting depth of the keyroot that owns node's right path
  const int32_t* kdn;    // keyroot nesting height at keyroots, -1 elsewhere
  const int32_t* lrank;  // leaves strictly before node (n+1 entries per tree)
  const int32_t* kr;     // keyroot node ids, descending
  const int32_t* krkd;   // nesting height per keyroot
  const int32_t* krsize; // subtree size per keyroot
  const int32_t* krleaf; // leaves in subtree per keyroot
  const int32_t* skr;    // sorted keyroots (leaf keyroots, then by nesting height)
  const int32_t* skd;
  const int32_t* ssize;
  const int32_t* sleaf;
  const int32_t* spre;   // prefix sums of ssize (K+1 entries)
  const int32_t* kdoff;
  const int32_t* rtop;   // keyroot at the top of node's right path
  const int32_t* pdel;   // cost mode: deletion costs of nodes [0, i) (n+1 entries per tree)
  const int32_t* pins;   // cost mode: insertion costs of nodes [0, i)
  const NodeRec* crec;   // per node, as a table column
  const NodeRec* rrec;   // per node, as a table row
  const CostRec* ccrec;  // cost mode: per node, pdel / pins at it and past its subtree
  const TreeMeta* trees;
};

// A table task: pair id (24 bits) | x keyroot node (20 bits) | y keyroot node (20 bits).
// Trees in a pool are limited to 2^20 nodes and a batch to 2^24 pairs.
constexpr int kTaskNodeBits = 20;
constexpr int kTaskPairBits = 24;
XTED_HD inline uint64_t task_pack(uint32_t pair, uint32_t i0, uint32_t j0) {
  return ((uint64_t)pair << 40) | ((uint64_t)i0 << 20) | (uint64_t)j0;
}
XTED_HD inline uint32_t task_pair(uint64_t t) { return (uint32_t)(t >> 40); }
XTED_HD inline uint32_t task_i0(uint64_t t) { return (uint32_t)(t >> 20) & 0xFFFFFu; }
XTED_HD inline uint32_t task_j0(uint64_t t) { return (uint32_t)t & 0xFFFFFu; }

// ---- granularity classes ----------------------------------------------------
// Class 0: leaf (a table whose short side is one node, a leaf keyroot),
//   evaluated in closed form by the leaf kernel, one line per leaf keyroot.
// Classes 1..25: workgroup row scan with W = 2^wi waves and C = 2^ci columns
//   per lane, class = 1 + wi*5 + ci (wi, ci in 0..4).
// Class 26: tables no form below can hold (an error).
// Classes 27..32: column-blocked form for tables too wide or too LDS-hungry
//   for one workgroup: blocks of xl_bw(k) columns, one workgroup each,
//   chained per row through device memory, class = 27 + k.
// Classes 33..35: tiny tables of at most 16, 32 or 64 columns, one per 16-,
//   32- or 64-lane segment of a wave (four, two or one to a wave).
constexpr int kClassLane = 0;  // the leaf class
constexpr int kNumRowClasses = 25;
constexpr int kClassXL = 26;
constexpr int kFirstXLClass = 27;
constexpr int kNumXLForms = 6;
constexpr int kClassTiny = kFirstXLClass + kNumXLForms;  // 33: the first tiny form
constexpr int kNumTinyForms = 3;
constexpr int kNumClasses = kClassTiny + kNumTinyForms;  // 36
XTED_HD inline bool is_tiny(int cls) { return cls >= kClassTiny && cls < kNumClasses; }
// Lanes per table (and most columns) of tiny form k.
XTED_HD inline int tiny_seg(int k) { return 16 << k; }
// Slot bytes of one table of tiny form k: 32 KiB per workgroup of 256 lanes.
XTED_HD inline int64_t tiny_max_lds(int k) { return (int64_t)2048 << k; }
// Column-blocked forms: waves and columns per lane of one block.
XTED_HD inline int xl_waves(int k) { return k <= 1 ? 8 : k == 2 ? 4 : 2; }
XTED_HD inline int xl_cpl(int k) { return k == 0 ? 8 : k <= 3 ? 4 : k == 4 ? 2 : 1; }
XTED_HD inline int xl_bw(int k) { return 64 * xl_waves(k) * xl_cpl(k); }  // 4096 .. 128
// Column-blocked tables with at most this many rows are swept block after
// block by one workgroup (forms of at most 1024 columns), with no waits
// between workgroups; taller ones pipeline one workgroup per block.
constexpr int32_t kXLSeqRows = 256;  // default of ClassRule::xl_seq_rows
constexpr int kFirstSeqXLForm = 2;
// Far values a block takes from earlier blocks are indexed by the nesting
// depth of their right path below the column keyroot: at most this many.
constexpr int kMaxImports = 256;
XTED_HD inline int row_class(int wi, int ci) { return 1 + wi * 5 + ci; }

// Each (level, class) bucket is cut into size bands, largest tables first,
// so that a launch starts its longest tables first and its last workgroups
// are short ones (column-blocked classes use band 0 only). The plan's
// buckets are (level, class, band), and one launch covers kRowBands
// consecutive bands.
//
// A launch gives every workgroup the largest slot storage of its tables,
// and LDS is what bounds how many workgroups a CU holds. In the 4096-pair
// batch a W = 2 launch's tables need 2 KiB of slot rows on average and its
// largest 28 KiB, which held a CU to 5 workgroups where registers allow 10.
// So with mode 2 the row-scan tables that need more than light_per_wave
// bytes per wave (8 KiB is one workgroup's share of LDS at 20 waves per CU)
// go to bands of their own (kRowBands .. 2*kRowBands-1), and a (level,
// class) takes two launches: light tables and heavy ones.
constexpr int kRowBands = 4;
constexpr int kBands = 2 * kRowBands;
constexpr int kLightPerWave = 10 * 1024;  // default of Context::light_per_wave
// mode 0: band 0 (plan order); 1: row bands; 2: row bands, light / heavy.
XTED_HD inline int size_band(int cls, int rows, int64_t lds, int mode, int light_per_wave) {
  if (mode == 0) return 0;
  if (is_tiny(cls)) {
    const int q = tiny_seg(cls - kClassTiny) / 4;  // row bands of a quarter segment
    return rows > 3 * q ? 0 : rows > 2 * q ? 1 : rows > q ? 2 : 3;
  }
  if (cls < 1 || cls > kNumRowClasses) return 0;
  const int rb = rows >= 256 ? 0 : rows >= 64 ? 1 : rows >= 16 ? 2 : 3;
  const int64_t light = (int64_t)(1 << ((cls - 1) / 5)) * light_per_wave;
  return mode == 2 && lds > light ? kRowBands + rb : rb;
}
XTED_HD inline int bucket_of(int level, int cls, int band) {
  return (level * kNumClasses + cls) * kBands + band;
}

// Saved-row slots are uint16 (F <= |subtree x| + |subtree y| < 65536).
// LDS budget for the slots of one table, per wave of its workgroup.
constexpr int kLdsPerWave = 10 * 1024;
constexpr int kLdsMaxTable = 150 * 1024;
constexpr int kMaxRowWaves = 16;
constexpr int kMaxCols = 64 * 16 * kMaxRowWaves;  // 16384
// Widest table the on-chip forms take (W = 16 waves x 8 columns per lane;
// 16 columns per lane at 16 waves would spill); wider tables are column-blocked.
constexpr int kMaxOnChipCols = 64 * 8 * kMaxRowWaves;  // 8192

// 16-lane form only by default. Under the default rule the 32-lane form ran
// the 4096-pair batch in the same time (100.6 against 100.9 ms unit, 125.0
// against 124.4 ms under a cost model); the tuned rule (x-ted_amd/tuning.json:
// up to 16 columns per lane) turns it on (96.5 against 99.6 ms). The 64-lane
// form ran 4% slower (4 waves per SIMD for its 64 TD registers per lane).
constexpr int32_t kTinyDefault = 1;
// Most columns per lane a table takes under a cost model.
constexpr int32_t kMaxCostCpl = 8;

// Tuning knobs carried by the plan builder (autotuner writes these).
struct ClassRule {
  int32_t lds_per_wave;  // bytes of slot storage one wave may own
  int32_t min_waves;     // lower bound on W for tables with >= wide_cols columns
  int32_t wide_cols;
  int32_t lds_max;       // slot bytes one workgroup may hold (device limit)
  int32_t max_cpl;       // most columns per lane before a table gets more waves
  int32_t xl_seq_rows = kXLSeqRows;  // column-blocked tables this short: one workgroup
  // Tiny forms in use (bit k: tables of at most tiny_seg(k) columns take
  // form k, the narrowest that holds them).
  int32_t tiny = kTinyDefault;
  // Widest form a pipelined column-blocked table may use (0: 4096-column
  // blocks, 1: 2048, 2: 1024, ...): narrower blocks put more workgroups on
  // one table.
  int32_t xl_first = 0;
};

XTED_HD inline int ilog2_ceil(int v) {
  int r = 0;
  while ((1 << r) < v) ++r;
  return r;
}

// Slot bytes of one block of a column-blocked table: one uint16 row per
// nesting level of the row keyroot, each holding the block's own far columns
// (at most min(leaves, bw)), the boundary entry, kd_c+1 imports and a sink.
XTED_HD inline int64_t xl_lds_need(int kd_r, int kd_c, int leaves_c, int bw) {
  const int own = leaves_c < bw ? leaves_c : bw;
  return (int64_t)(kd_r > 1 ? kd_r : 1) * (own + 3 + kd_c) * 2;
}

// Class of table (a x b): lds_need is the slot storage of its oriented
// on-chip form; kd_r / kd_c / leaves_c (row keyroot height, column keyroot
// height and leaves) size the column-blocked forms. Single-row tables are
// the leaf class when leaf_form is set. *need receives the slot bytes of the
// chosen form.
XTED_HD inline int table_class(int a, int b, int64_t lds_need, int kd_r, int kd_c, int leaves_c,
                               bool leaf_form, const ClassRule& rule, int64_t* need) {
  const int rows = a <= b ? a : b;
  const int cols = a <= b ? b : a;
  *need = lds_need;
  if (rows == 1 && leaf_form) return kClassLane;
  const int cpl_cap = rule.max_cpl < 8 ? rule.max_cpl : 8;
  if (cols > 64 * cpl_cap * kMaxRowWaves || lds_need > rule.lds_max) {
    if (kd_c + 1 > kMaxImports) return kClassXL;
    // Short tables take the forms one workgroup sweeps.
    for (int k = rows <= rule.xl_seq_rows ? kFirstSeqXLForm : rule.xl_first; k < kNumXLForms;
         ++k) {
      const int64_t n = xl_lds_need(kd_r, kd_c, leaves_c, xl_bw(k));
      if (n <= rule.lds_max) {
        *need = n;
        return kFirstXLClass + k;
      }
    }
    return kClassXL;
  }
  for (int k = 0; k < kNumTinyForms; ++k)
    if ((rule.tiny >> k & 1) && cols <= tiny_seg(k) && lds_need <= tiny_max_lds(k))
      return kClassTiny + k;
  int wi = 0;
  while (wi < 4 && lds_need > (int64_t)rule.lds_per_wave * (1 << wi)) ++wi;
  while (wi < 4 && cols > 64 * rule.max_cpl * (1 << wi)) ++wi;
  if (cols >= rule.wide_cols) {
    while (wi < 4 && (1 << wi) < rule.min_waves) ++wi;
  }
  int w = 1 << wi;
  int c = (cols + 64 * w - 1) / (64 * w);
  int ci = ilog2_ceil(c);
  return row_class(wi, ci);
}

}  // namespace xted
