// Device-side level scheduler: every keyroot-pair table of every pair in a
// batch is assigned level = depth_x(k) + depth_y(l) and a granularity class,
// then scattered into (level, class) buckets, all on the device.
//
// Replaces the reference's per-level host loop (X-TED_GPU/parallel_X_TED.cu:
// 214-220 level array, :304-385 fetch_task / fetch_size_queue /
// thrust::sort_by_key / filter_new with several host syncs per level, and
// X-TED_GPU/function.cu:18-57). Here the whole plan is built once per batch
// by three kernels (count, scan, scatter) and read back with one copy of the
// bucket counts. Counts and appends are aggregated per wave and then per
// workgroup: each distinct bucket of a workgroup costs one global atomic, and
// a wave's tables of one bucket land contiguously.
#include <hip/hip_runtime.h>
#include <algorithm>

#include "kernels.hpp"

namespace xted {

constexpr int kPlanThreads = 256;

__device__ __forceinline__ int find_pair(const int64_t* blk_start, int npairs, int64_t b) {
  int lo = 0, hi = npairs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (blk_start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Per-table classification shared by both passes.
__device__ __forceinline__ void classify(const PlanArgs& A, const TreeMeta& tx,
                                         const TreeMeta& ty, int32_t loff, bool leaf_form, int k,
                                         int l, int* bucket,
                                         int* lds, uint32_t* i0, uint32_t* j0,
                                         unsigned long long* cells) {
  const int64_t kx = tx.node_off + k, ly = ty.node_off + l;
  const int32_t a = A.pool.ssize[kx], b = A.pool.ssize[ly];
  *cells = (unsigned long long)a * (unsigned long long)b;
  const int32_t kdx = A.pool.skd[kx], kdy = A.pool.skd[ly];
  const int32_t level = kdx + kdy + loff;
  const ClassRule& rule = A.tail_level[level] ? A.rule_tail : A.rule;
  // Slot storage of the oriented table: one uint16 row per nesting level of
  // the row keyroot (at least one: the kernels read slot 0 unconditionally),
  // each as long as the column keyroot's leaf count plus one sink entry.
  const int32_t slots = max(a <= b ? kdx : kdy, 1);
  const int32_t leaves_c = a <= b ? A.pool.sleaf[ly] : A.pool.sleaf[kx];
  int64_t need = (int64_t)slots * (leaves_c + 1) * 2;
  // With the leaf kernel off, single-row tables take the row-scan forms.
  int cls = table_class(a, b, need, a <= b ? kdx : kdy, a <= b ? kdy : kdx, leaves_c,
                        leaf_form, rule, &need);
  // A forced column-blocked form takes every multi-row table it can hold.
  if (A.force_class >= kFirstXLClass && A.force_class < kFirstXLClass + kNumXLForms && cls >= 1 &&
      cls != kClassXL) {
    const int k = A.force_class - kFirstXLClass;
    const int32_t kd_c = a <= b ? kdy : kdx;
    const int64_t n = xl_lds_need(a <= b ? kdx : kdy, kd_c, leaves_c, xl_bw(k));
    if (kd_c + 1 <= kMaxImports && n <= A.rule.lds_max) {
      cls = A.force_class;
      need = n;
    }
  }
  // A forced row-scan form (or tiny form) takes every on-chip table it can
  // hold.
  if (is_tiny(A.force_class) && cls >= 1 && (cls <= kNumRowClasses || is_tiny(cls))) {
    const int k = A.force_class - kClassTiny;
    const int cols = a <= b ? b : a;
    const int64_t row_need = (int64_t)slots * (leaves_c + 1) * 2;
    if (cols <= tiny_seg(k) && row_need <= tiny_max_lds(k)) {
      cls = A.force_class;
      need = row_need;
    }
  }
  if (A.force_class >= 1 && A.force_class <= kNumRowClasses && cls >= 1 &&
      (cls <= kNumRowClasses || is_tiny(cls))) {
    const int wi = (A.force_class - 1) / 5, ci = (A.force_class - 1) % 5;
    const int cols = a <= b ? b : a;
    const int64_t row_need = (int64_t)slots * (leaves_c + 1) * 2;
    if ((cls < kClassXL || is_tiny(cls)) && cols <= 64 * (1 << wi) * (1 << ci) &&
        row_need <= A.rule.lds_max) {
      cls = A.force_class;
      need = row_need;
    }
  }
  *bucket = bucket_of(level, cls, size_band(cls, a <= b ? a : b, need, A.bands, A.light_per_wave));
  *lds = (int)need;
  *i0 = (uint32_t)A.pool.skr[kx];
  *j0 = (uint32_t)A.pool.skr[ly];
}

#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

// Wave-wide sum / max of 32-bit values, result in every lane (DPP prefix,
// then lane 63 broadcast).
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, DPP_ROW_SHR(1), 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, DPP_ROW_SHR(2), 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, DPP_ROW_SHR(4), 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, DPP_ROW_SHR(8), 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, DPP_ROW_BCAST15, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, DPP_ROW_BCAST31, 0xc, 0xf, false);
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(1), 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(2), 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(4), 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(8), 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST15, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST31, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}

// Per-workgroup bucket table in LDS: the plan's global atomics go to a few
// thousand (level, class) buckets and device-scope atomics are performed
// memory-side, so a wave's per-bucket partials are first combined per
// workgroup (open addressing on the bucket id), and each workgroup then
// makes one global atomic per bucket it touched. A bucket that finds the
// table full takes the global atomics directly.
constexpr int kHTBits = 8;
constexpr int kHT = 1 << kHTBits;
struct PlanLds {
  int32_t key[kHT];  // bucket id, or -1
  uint32_t cnt[kHT];
  int32_t lds[kHT];
  unsigned long long cells[kHT];  // count pass: cells; scatter pass: base
};

__device__ __forceinline__ int ht_slot(PlanLds& T, int bucket, bool insert) {
  int h = (int)(((uint32_t)bucket * 0x9E3779B1u) >> (32 - kHTBits));
  for (int probe = 0; probe < kHT; ++probe) {
    const int k = insert ? atomicCAS(&T.key[h], -1, bucket) : T.key[h];
    if (k == bucket || (insert && k == -1)) return h;
    if (k == -1) return -1;  // lookup: absent
    h = (h + 1) & (kHT - 1);
  }
  return -1;  // full
}

__device__ __forceinline__ void ht_clear(PlanLds& T) {
  for (int i = threadIdx.x; i < kHT; i += kPlanThreads) {
    T.key[i] = -1;
    T.cnt[i] = 0;
    T.lds[i] = 0;
    T.cells[i] = 0;
  }
}

// The tables of this workgroup: y keyroots [l_lo, l_lo + rows) of the
// pair's sorted list against all Kx x keyroots, total of them.
struct PlanChunk {
  int p, kbase, lbase, Kx, l_lo;
  int64_t total;
};
__device__ __forceinline__ PlanChunk plan_chunk(const PlanArgs& A, const TreeMeta& tx,
                                                const TreeMeta& ty, int p) {
  PlanChunk c;
  c.p = p;
  // With the leaf kernel on, tables with a leaf keyroot are not planned here
  // (the leaf kernel writes them): only the non-leaf keyroots [K1, K) of the
  // sorted lists take part.
  const bool le = A.pair_leaf ? A.pair_leaf[p] != 0 : A.lane_enable != 0;
  c.kbase = le ? tx.K1 : 0;
  c.lbase = le ? ty.K1 : 0;
  c.Kx = tx.K - c.kbase;
  const int Ky = ty.K - c.lbase;
  const int lper = max(1, A.lchunk_tables / max(c.Kx, 1));  // same as plan_blocks()
  c.l_lo = (int)(blockIdx.x - A.blk_start[p]) * lper;
  const int l_hi = min(Ky, c.l_lo + lper);
  c.total = (int64_t)max(0, l_hi - c.l_lo) * c.Kx;
  return c;
}

// Count pass: per distinct bucket of a wave, the table count, the interior
// cells (a table has < 2^30 cells; the wave sum is taken over 16-bit halves
// so 32-bit sums cannot overflow) and the largest slot-storage need, added
// into the workgroup's table.
__device__ __forceinline__ void wave_count(const PlanArgs& A, PlanLds& T, int bucket, bool active,
                                           uint32_t cells, int32_t lds) {
  unsigned long long todo = __ballot(active);
  const int lane = (int)(threadIdx.x & 63);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lb = __shfl(bucket, leader);
    const bool mine = active && bucket == lb;
    const unsigned long long same = __ballot(mine);
    const uint32_t lo = wave_sum_u32(mine ? (cells & 0xFFFFu) : 0u);
    const uint32_t hi = wave_sum_u32(mine ? (cells >> 16) : 0u);
    const int32_t mx = wave_max_i32(mine ? lds : 0);
    if (lane == leader) {
      const unsigned long long n = (unsigned long long)__popcll(same);
      const unsigned long long sum = ((unsigned long long)hi << 16) + lo;
      const int h = ht_slot(T, lb, true);
      if (h >= 0) {
        atomicAdd(&T.cnt[h], (uint32_t)n);
        atomicAdd(&T.cells[h], sum);
        atomicMax(&T.lds[h], mx);
      } else {
        atomicAdd(&A.counts[lb], n);
        atomicAdd(&A.cells[lb], sum);
        if (mx > 0) atomicMax(&A.maxlds[lb], mx);
      }
    }
    todo &= ~same;
  }
}

// Scatter pass, first sweep: tables per distinct bucket into the table.
__device__ __forceinline__ void wave_tally(PlanLds& T, int bucket, bool active) {
  unsigned long long todo = __ballot(active);
  const int lane = (int)(threadIdx.x & 63);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lb = __shfl(bucket, leader);
    const unsigned long long same = __ballot(active && bucket == lb);
    if (lane == leader) {
      const int h = ht_slot(T, lb, true);
      if (h >= 0) atomicAdd(&T.cnt[h], (uint32_t)__popcll(same));
    }
    todo &= ~same;
  }
}

// Scatter pass, second sweep: this lane's slot, from the workgroup's range
// of the bucket (or, for a bucket not in the table, a global reservation).
__device__ __forceinline__ unsigned long long wave_reserve(const PlanArgs& A, PlanLds& T,
                                                           int bucket, bool active) {
  unsigned long long pos = 0;
  const unsigned long long lt = __lanemask_lt();
  unsigned long long todo = __ballot(active);
  const int lane = (int)(threadIdx.x & 63);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lb = __shfl(bucket, leader);
    const bool mine = active && bucket == lb;
    const unsigned long long same = __ballot(mine);
    unsigned long long base = 0;
    if (lane == leader) {
      const uint32_t n = (uint32_t)__popcll(same);
      const int h = ht_slot(T, lb, false);
      base = h >= 0 ? T.cells[h] + atomicAdd(&T.cnt[h], n)
                    : atomicAdd(&A.cursor[lb], (unsigned long long)n);
    }
    base = __shfl(base, leader);
    if (mine) pos = base + __popcll(same & lt);
    todo &= ~same;
  }
  return pos;
}

// The start-up values of table (i0, j0) of a pair.
__device__ __forceinline__ TaskDesc make_desc(const PlanArgs& A, const PairMeta& pm,
                                              const TreeMeta& tx, const TreeMeta& ty, int32_t i0,
                                              int32_t j0) {
  const int32_t a = A.pool.orl[tx.node_off + i0] - i0 + 1;
  const int32_t b = A.pool.orl[ty.node_off + j0] - j0 + 1;
  const bool rows_x = a <= b;
  TaskDesc d;
  d.tdo = rows_x ? pm.td_off : pm.tdt_off;
  d.tdt = rows_x ? pm.tdt_off : pm.td_off;
  d.offR = rows_x ? tx.node_off : ty.node_off;
  d.offC = rows_x ? ty.node_off : tx.node_off;
  d.ldo = rows_x ? ty.n : tx.n;
  d.ldt = rows_x ? tx.n : ty.n;
  d.r0 = rows_x ? i0 : j0;
  d.c0 = rows_x ? j0 : i0;
  d.rmax = d.r0 + (rows_x ? a : b);
  d.cmax = d.c0 + (rows_x ? b : a);
  d.base_rx = A.pool.kdep[d.offR + d.r0] | (rows_x ? (int32_t)0x80000000 : 0);
  d.lbase = A.pool.lrank[d.offC + d.c0];
  return d;
}

template <bool kScatter>
__global__ __launch_bounds__(kPlanThreads) void plan_kernel(PlanArgs A) {
  __shared__ PlanLds T;
  const int p = find_pair(A.blk_start, A.npairs, blockIdx.x);
  const PairMeta pm = A.pairs[p];
  const TreeMeta tx = A.pool.trees[pm.tx], ty = A.pool.trees[pm.ty];
  const PlanChunk ch = plan_chunk(A, tx, ty, p);
  const int32_t loff = A.level_off ? A.level_off[p] : 0;
  const bool leaf_form = A.pair_leaf ? A.pair_leaf[p] != 0 : A.lane_enable != 0;
  ht_clear(T);
  __syncthreads();
  // Table idx of the chunk is (l, k) = (idx / Kx, idx % Kx); a chunk holds
  // at most max(lchunk_tables, Kx) < 2^15 tables. One division per thread,
  // then each step of kPlanThreads tables adds (dl, dk) with one carry.
  auto sweep = [&](auto body) {
    const int32_t total = (int32_t)ch.total;
    const int32_t kx = max(ch.Kx, 1);
    const int32_t dl = kPlanThreads / kx, dk = kPlanThreads % kx;
    int32_t lq = (int32_t)threadIdx.x / kx, kr = (int32_t)threadIdx.x % kx;
    for (int32_t base = 0; base < total; base += kPlanThreads) {
      const bool active = base + (int32_t)threadIdx.x < total;
      int bucket = 0, lds = 0;
      uint32_t i0 = 0, j0 = 0;
      unsigned long long cells = 0;
      if (active)
        classify(A, tx, ty, loff, leaf_form, ch.kbase + kr, ch.lbase + ch.l_lo + lq, &bucket,
                 &lds, &i0, &j0, &cells);
      body(active, bucket, lds, i0, j0, cells);
      lq += dl;
      kr += dk;
      if (kr >= kx) {
        kr -= kx;
        ++lq;
      }
    }
  };
  if (!kScatter) {
    sweep([&](bool active, int bucket, int lds, uint32_t, uint32_t, unsigned long long cells) {
      wave_count(A, T, bucket, active, (uint32_t)cells, lds);
    });
    __syncthreads();
    for (int h = threadIdx.x; h < kHT; h += kPlanThreads) {
      const int b = T.key[h];
      if (b < 0) continue;
      atomicAdd(&A.counts[b], (unsigned long long)T.cnt[h]);
      atomicAdd(&A.cells[b], T.cells[h]);
      if (T.lds[h] > 0) atomicMax(&A.maxlds[b], T.lds[h]);
    }
  } else {
    sweep([&](bool active, int bucket, int, uint32_t, uint32_t, unsigned long long) {
      wave_tally(T, bucket, active);
    });
    __syncthreads();
    // One range per bucket for the workgroup; cnt becomes its cursor.
    for (int h = threadIdx.x; h < kHT; h += kPlanThreads) {
      const int b = T.key[h];
      if (b < 0) continue;
      T.cells[h] = atomicAdd(&A.cursor[b], (unsigned long long)T.cnt[h]);
      T.cnt[h] = 0;
    }
    __syncthreads();
    sweep([&](bool active, int bucket, int, uint32_t i0, uint32_t j0, unsigned long long) {
      const unsigned long long pos = wave_reserve(A, T, bucket, active);
      if (active) {
        A.tasks[pos] = task_pack((uint32_t)ch.p, i0, j0);
        A.descs[pos] = make_desc(A, pm, tx, ty, (int32_t)i0, (int32_t)j0);
      }
    });
  }
}

// Exclusive scan of the bucket counts into the scatter cursors (one block).
__global__ __launch_bounds__(1024) void plan_scan_kernel(const unsigned long long* counts,
                                                          unsigned long long* cursor, int nb) {
  __shared__ unsigned long long part[1024];
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int lo = min(nb, tid * per), hi = min(nb, lo + per);
  unsigned long long s = 0;
  for (int i = lo; i < hi; ++i) s += counts[i];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    unsigned long long v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  unsigned long long run = tid ? part[tid - 1] : 0;
  for (int i = lo; i < hi; ++i) {
    cursor[i] = run;
    run += counts[i];
  }
}

// Clears the count / max / ticket arrays and the abort flag (a kernel rather
// than a memset so the plan can be captured into a graph).
__global__ void plan_clear_kernel(PlanArgs a, int nb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) {
    a.counts[i] = 0;
    a.cells[i] = 0;
    a.maxlds[i] = 0;
    a.tickets[i] = 0;
  }
  if (i < kAbortInts) a.abort_flag[i] = 0;
}

void launch_plan(const PlanArgs& a, int64_t nblocks, hipStream_t s) {
  const int nb = a.nlevels * kNumClasses * kBands;
  hipLaunchKernelGGL(plan_clear_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, a, nb);
  if (nblocks > 0)
    hipLaunchKernelGGL(plan_kernel<false>, dim3((unsigned)nblocks), dim3(kPlanThreads), 0, s, a);
  hipLaunchKernelGGL(plan_scan_kernel, dim3(1), dim3(1024), 0, s, a.counts, a.cursor, nb);
  if (nblocks > 0)
    hipLaunchKernelGGL(plan_kernel<true>, dim3((unsigned)nblocks), dim3(kPlanThreads), 0, s, a);
}

// Distances TD[0][0] of every pair into dist[].
__global__ void gather_dist_kernel(const PairMeta* pairs, const td_t* td, int32_t npairs,
                                   int32_t* dist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npairs) dist[i] = td[pairs[i].td_off];
}

void launch_gather_dist(const PairMeta* pairs, const td_t* td, int32_t npairs, int32_t* dist,
                        hipStream_t s) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(gather_dist_kernel, dim3((npairs + 255) / 256), dim3(256), 0, s, pairs, td,
                     npairs, dist);
}

__global__ __launch_bounds__(256) void widen_td_kernel(const td_t* __restrict__ src,
                                                       int32_t* __restrict__ dst, int64_t n) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
    dst[i] = src[i];
}

void launch_widen_td(const td_t* src, int32_t* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(widen_td_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
}

}  // namespace xted
