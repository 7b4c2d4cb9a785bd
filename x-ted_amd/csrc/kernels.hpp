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
  int32_t* maxlds;             // kNumClasses
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
// Launches the kernel of class `cls` over a.tasks[0..ntasks). Returns false
// when the class has no kernel.
bool launch_forest(int cls, const DPArgs& a, int lds_bytes, hipStream_t s);

}  // namespace xted
