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
  TreeMeta* trees;
  int32_t* scratch;            // 6*(n+1) ints per tree at scratch_off[b]
  const int64_t* scratch_off;  // per block of this launch
  int32_t first_tree;
};
void launch_tree_prep(const PrepArgs& a, int ntrees, hipStream_t s);

// Plan builder: count, scan and scatter of (level, class) buckets.
struct PlanArgs {
  PoolView pool;
  const PairMeta* pairs;
  const int64_t* blk_start;    // npairs+1: first block of each pair
  int32_t npairs;
  int32_t nlevels;
  int32_t lchunk_tables;       // target tables per block
  ClassRule rule;
  int32_t force_class;         // -1 = automatic; else row-scan tables go to this class if they fit
  int32_t lane_enable;         // 0 = single-row tables use the row-scan kernels
  unsigned long long* counts;  // nlevels*kNumClasses
  unsigned long long* cells;   // nlevels*kNumClasses: interior cells per bucket
  unsigned long long* cursor;  // scatter cursors (bucket offsets)
  int32_t* maxlds;             // nlevels*kNumClasses: largest slot storage per bucket
  uint64_t* tasks;
};
// Blocks a pair occupies in the plan kernels (same formula on both sides).
inline int64_t plan_blocks(int32_t Kx, int32_t Ky, int32_t chunk) {
  const int32_t lper = chunk / (Kx > 0 ? Kx : 1) > 0 ? chunk / (Kx > 0 ? Kx : 1) : 1;
  return (Ky + lper - 1) / lper;
}
void launch_plan(const PlanArgs& a, int64_t nblocks, hipStream_t s);
void launch_gather_dist(const PairMeta* pairs, const int32_t* td, int32_t npairs, int32_t* dist,
                        hipStream_t s);

// Tables with a leaf keyroot, enumerated per level without a task list.
struct LaneArgs {
  PoolView pool;
  const PairMeta* pairs;
  int32_t npairs;
  int32_t level;                  // level of a lane launch
  unsigned long long* offsets;    // nlevels*(npairs+1) prefix of rectangle sizes
  unsigned long long* totals;     // nlevels: tables
  unsigned long long* cells;      // nlevels: interior cells
  int32_t* td;
};

// |A| + |B| of pair p at level L (A: leaf keyroots of x by keyroots of y of
// height L; B: non-leaf keyroots of x of height L by leaf keyroots of y).
// When sizes is given it receives {|A|, |Y_L|, |X'_L|, |B|} and the start
// positions {ylo, xlo} of Y_L and X'_L in the sorted lists.
XTED_HD inline unsigned long long lane_rect_count(const LaneArgs& A, int p, int L, int32_t* sizes) {
  const PairMeta pm = A.pairs[p];
  const TreeMeta tx = A.pool.trees[pm.tx], ty = A.pool.trees[pm.ty];
  const int32_t nly = L <= ty.kdmax ? A.pool.kdoff[ty.node_off + L + 1] - A.pool.kdoff[ty.node_off + L] : 0;
  const int32_t nlx = L <= tx.kdmax ? A.pool.kdoff[tx.node_off + L + 1] - A.pool.kdoff[tx.node_off + L] : 0;
  const int32_t ny = (L == 0 ? ty.K1 : 0) + nly;
  const int32_t ylo = L == 0 ? 0 : ty.K1 + (L <= ty.kdmax ? A.pool.kdoff[ty.node_off + L] : 0);
  const int32_t xlo = tx.K1 + (L <= tx.kdmax ? A.pool.kdoff[tx.node_off + L] : 0);
  const unsigned long long ca = (unsigned long long)tx.K1 * (unsigned long long)ny;
  const unsigned long long cb = (unsigned long long)nlx * (unsigned long long)ty.K1;
  if (sizes) {
    sizes[0] = (int32_t)ca;
    sizes[1] = ny;
    sizes[2] = nlx;
    sizes[3] = ylo;
    sizes[4] = xlo;
  }
  return ca + cb;
}
// Interior cells of pair p's rectangles at level L (sizes from lane_rect_count).
XTED_HD inline unsigned long long lane_rect_cells(const LaneArgs& A, int p, const int32_t* sz) {
  const PairMeta pm = A.pairs[p];
  const TreeMeta tx = A.pool.trees[pm.tx], ty = A.pool.trees[pm.ty];
  const int32_t* py = A.pool.spre + ty.node_off;
  const int32_t* px = A.pool.spre + tx.node_off;
  return (unsigned long long)tx.K1 * (unsigned long long)(py[sz[3] + sz[1]] - py[sz[3]]) +
         (unsigned long long)ty.K1 * (unsigned long long)(px[sz[4] + sz[2]] - px[sz[4]]);
}
void launch_lane_offsets(const LaneArgs& a, int nlevels, hipStream_t s);
void launch_lane_level(const LaneArgs& a, unsigned long long total, hipStream_t s);

// Ints of padding the TD arena keeps before and after every table, so the
// row-scan kernels can issue whole-group loads without bounds branches.
constexpr int kTdPad = 4;

// Forest DP.
struct DPArgs {
  PoolView pool;
  const PairMeta* pairs;
  int32_t* td;            // arena holding every TD / TDT
  const uint64_t* tasks;  // this launch's tables
  int64_t ntasks;
};
// Lets every row-scan kernel request up to lds_limit bytes of dynamic LDS.
hipError_t forest_init_attributes(int lds_limit);
// Launches the row-scan kernel of class `cls` over a.tasks[0..ntasks).
// Returns false when the class has no kernel (the lane class runs per level
// through launch_lane_level, the XL class has none).
bool launch_forest(int cls, const DPArgs& a, int lds_bytes, hipStream_t s);

}  // namespace xted
