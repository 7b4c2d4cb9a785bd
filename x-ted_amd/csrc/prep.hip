// Device-side tree preprocessing: outermost-right leaves (orl), keyroots,
// keyroot nesting heights (the level scheduler's depths), per-node path
// keyroot depth, leaf ranks and the descending keyroot list.
//
// Replaces the host loops of X-TED_GPU/TED.cu:21-75 (outermost_right_leaves,
// key_roots) and X-TED_GPU/parallel_X_TED.cu:184-212 (O(K^2) keyroot depth).
// One workgroup per tree; every phase is a data-parallel pass separated by
// workgroup barriers, with pointer jumping for the ancestor-chain quantities.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace xted {

constexpr int kPrepThreads = 1024;

// Block-wide exclusive prefix sum over flag[0..n) into out[0..n]; out[n] = total.
__device__ static void block_exclusive_scan(const int32_t* flag, int32_t* out, int32_t n,
                                            int32_t* lds_tmp) {
  const int tid = threadIdx.x;
  // Each thread owns a contiguous chunk.
  const int per = (n + kPrepThreads - 1) / kPrepThreads;
  const int lo = min(n, tid * per), hi = min(n, lo + per);
  int32_t s = 0;
  for (int i = lo; i < hi; ++i) s += flag[i];
  lds_tmp[tid] = s;
  __syncthreads();
  // Hillis-Steele over 1024 partial sums.
  for (int off = 1; off < kPrepThreads; off <<= 1) {
    int32_t v = tid >= off ? lds_tmp[tid - off] : 0;
    __syncthreads();
    lds_tmp[tid] += v;
    __syncthreads();
  }
  int32_t run = tid ? lds_tmp[tid - 1] : 0;
  for (int i = lo; i < hi; ++i) {
    out[i] = run;
    run += flag[i];
  }
  if (tid == kPrepThreads - 1) out[n] = lds_tmp[kPrepThreads - 1];
  __syncthreads();
}

__global__ __launch_bounds__(kPrepThreads) void tree_prep_kernel(PrepArgs A) {
  __shared__ int32_t lds_tmp[kPrepThreads];
  __shared__ int32_t changed;
  const int t = A.first_tree + blockIdx.x;
  TreeMeta meta = A.trees[t];
  const int32_t n = meta.n;
  const int64_t off = meta.node_off;
  const int32_t* parent = A.parent + off;
  const int32_t* label = A.label + off;
  (void)label;
  int32_t* orl = A.orl + off;
  int32_t* kdep = A.kdep + off;
  int32_t* kdn = A.kdn + off;
  int32_t* lrank = A.lrank + off;
  int32_t* S = A.scratch + A.scratch_off[blockIdx.x];
  int32_t* ptr = S;                  // last-child chains -> orl ; later path tops
  int32_t* anc0 = S + (n + 1);       // ancestor pointer (jumping), buffer 0
  int32_t* cnt0 = S + 2 * (n + 1);   // keyroot count on the jumped segment, buffer 0
  int32_t* anc1 = S + 3 * (n + 1);
  int32_t* cnt1 = S + 4 * (n + 1);
  int32_t* flag = S + 5 * (n + 1);   // keyroot / leaf flags
  const int tid = threadIdx.x;

  // ---- orl: follow last children to the outermost right leaf ---------------
  // In preorder the last child of v is its largest child id.
  for (int i = tid; i < n; i += kPrepThreads) ptr[i] = i;
  __syncthreads();
  for (int i = tid + 1; i < n; i += kPrepThreads) atomicMax(&ptr[parent[i]], i);
  __syncthreads();
  for (;;) {
    if (tid == 0) changed = 0;
    __syncthreads();
    int any = 0;
    for (int i = tid; i < n; i += kPrepThreads) {
      int32_t p = ptr[i], q = ptr[p];
      if (q != p) {
        ptr[i] = q;
        any = 1;
      }
    }
    if (any) changed = 1;
    __syncthreads();
    if (!changed) break;
    __syncthreads();
  }
  for (int i = tid; i < n; i += kPrepThreads) orl[i] = ptr[i];
  __syncthreads();

  // ---- keyroot flags: the root and every node that is not a last child ----
  for (int i = tid; i < n; i += kPrepThreads)
    flag[i] = (i == 0 || orl[parent[i]] != orl[i]) ? 1 : 0;
  __syncthreads();

  // ---- kanc(v): keyroots among the proper ancestors of v ---------------------
  // Pointer jumping with segment sums, double buffered.
  for (int i = tid; i < n; i += kPrepThreads) {
    anc0[i] = i ? parent[i] : -1;
    cnt0[i] = i ? flag[parent[i]] : 0;
  }
  __syncthreads();
  int32_t *ac = anc0, *cc = cnt0, *an = anc1, *cn = cnt1;
  for (;;) {
    if (tid == 0) changed = 0;
    __syncthreads();
    int any = 0;
    for (int i = tid; i < n; i += kPrepThreads) {
      int32_t a = ac[i];
      if (a >= 0) {
        an[i] = ac[a];
        cn[i] = cc[i] + cc[a];
        any = 1;
      } else {
        an[i] = -1;
        cn[i] = cc[i];
      }
    }
    if (any) changed = 1;
    __syncthreads();
    int32_t* tmp = ac; ac = an; an = tmp;
    tmp = cc; cc = cn; cn = tmp;
    if (!changed) break;
    __syncthreads();
  }
  // cc[i] now holds kanc(i).

  // ---- path top of v (keyroot owning v's right path) and kdep ---------------
  for (int i = tid; i < n; i += kPrepThreads) ptr[i] = flag[i] ? i : parent[i];
  __syncthreads();
  for (;;) {
    if (tid == 0) changed = 0;
    __syncthreads();
    int any = 0;
    for (int i = tid; i < n; i += kPrepThreads) {
      int32_t p = ptr[i], q = ptr[p];
      if (q != p) {
        ptr[i] = q;
        any = 1;
      }
    }
    if (any) changed = 1;
    __syncthreads();
    if (!changed) break;
    __syncthreads();
  }
  for (int i = tid; i < n; i += kPrepThreads) {
    kdep[i] = cc[ptr[i]];
    A.rtop[off + i] = ptr[i];
  }
  __syncthreads();

  // ---- keyroot nesting height ------------------------------------------------
  // kd(t) = max over keyroots u in (t, orl(t)] of kanc(u) - kanc(t), 0 if none.
  // Block maxima of kanc over keyroots (64 nodes per block) speed up the scans.
  int32_t* bmax = an;  // reuse a free buffer: (n+63)/64 entries
  const int nb = (n + 63) / 64;
  for (int b = tid; b < nb; b += kPrepThreads) {
    int32_t m = -1;
    for (int i = b * 64; i < min(n, b * 64 + 64); ++i)
      if (flag[i]) m = max(m, cc[i]);
    bmax[b] = m;
  }
  __syncthreads();
  for (int i = tid; i < n; i += kPrepThreads) {
    if (!flag[i]) {
      kdn[i] = -1;
      continue;
    }
    const int lo = i + 1, hi = orl[i];  // inclusive range
    int32_t m = -1;
    int j = lo;
    for (; j <= hi && (j & 63); ++j)
      if (flag[j]) m = max(m, cc[j]);
    for (; j + 63 <= hi; j += 64) m = max(m, bmax[j >> 6]);
    for (; j <= hi; ++j)
      if (flag[j]) m = max(m, cc[j]);
    kdn[i] = m < 0 ? 0 : m - cc[i];
  }
  __syncthreads();

  // ---- keyroot list, descending preorder -------------------------------------
  // cn is free now: exclusive scan of keyroot flags.
  block_exclusive_scan(flag, cn, n, lds_tmp);
  const int32_t K = cn[n];
  for (int i = tid; i < n; i += kPrepThreads) {
    if (flag[i]) {
      int32_t k = K - 1 - cn[i];
      A.kr[off + k] = i;
      A.krkd[off + k] = kdn[i];
      A.krsize[off + k] = orl[i] - i + 1;
    }
  }
  __syncthreads();

  // ---- leaf ranks ----------------------------------------------------------
  for (int i = tid; i < n; i += kPrepThreads) flag[i] = (orl[i] == i) ? 1 : 0;
  __syncthreads();
  block_exclusive_scan(flag, lrank, n, lds_tmp);
  for (int k = tid; k < K; k += kPrepThreads) {
    int32_t i = A.kr[off + k];
    A.krleaf[off + k] = lrank[orl[i] + 1] - lrank[i];
  }
  __syncthreads();

  // ---- per-node records of the forest kernels ------------------------------
  for (int i = tid; i < n; i += kPrepThreads) {
    const int32_t prev_leaf = i > 0 && orl[i - 1] == i - 1 ? (int32_t)0x80000000 : 0;
    A.crec[off + i] = NodeRec{orl[i], label[i], lrank[orl[i]], (i > 0 ? lrank[i - 1] : 0) | prev_leaf};
    A.rrec[off + i] = NodeRec{orl[i], label[i], kdep[i], (i > 0 ? kdep[i - 1] : 0) | prev_leaf};
  }
  __syncthreads();

  // ---- keyroots sorted for the scheduler -----------------------------------
  // Leaf keyroots first, then the other keyroots by ascending nesting height
  // and, within a height, by subtree size band (<= 16, 64, 256, 1024, more
  // nodes): a stable counting sort over the descending list. The plan gives
  // a wave consecutive keyroots of this order, so its tables fall into few
  // (level, class) buckets. One thread: this runs once per tree at upload.
  if (tid == 0) {
    constexpr int kBands = 5;
    auto band = [](int32_t sz) { return sz <= 16 ? 0 : sz <= 64 ? 1 : sz <= 256 ? 2 : sz <= 1024 ? 3 : 4; };
    const int32_t kdmax = kdn[0];
    // (kdmax+1)*kBands <= 5(n+1) entries: anc0 .. flag are free here.
    int32_t* hist = anc0;
    for (int d = 0; d < (kdmax + 1) * kBands; ++d) hist[d] = 0;
    int32_t K1 = 0;
    for (int k = 0; k < K; ++k) {
      const int32_t i = A.kr[off + k];
      if (orl[i] == i) ++K1;
      else ++hist[kdn[i] * kBands + band(orl[i] - i + 1)];
    }
    int32_t run = 0;
    for (int d = 0; d <= kdmax; ++d) {
      A.kdoff[off + d] = run;
      for (int b = 0; b < kBands; ++b) {
        const int32_t c = hist[d * kBands + b];
        hist[d * kBands + b] = run;
        run += c;
      }
    }
    A.kdoff[off + kdmax + 1] = run;
    int32_t leafpos = 0;
    for (int k = 0; k < K; ++k) {
      const int32_t i = A.kr[off + k];
      const int32_t pos =
          (orl[i] == i) ? leafpos++ : K1 + hist[kdn[i] * kBands + band(orl[i] - i + 1)]++;
      A.skr[off + pos] = i;
      A.skd[off + pos] = kdn[i];
      A.ssize[off + pos] = orl[i] - i + 1;
      A.sleaf[off + pos] = lrank[orl[i] + 1] - lrank[i];
    }
    int32_t acc = 0;
    for (int k = 0; k < K; ++k) {
      A.spre[off + k] = acc;
      acc += A.ssize[off + k];
    }
    A.spre[off + K] = acc;
    A.trees[t].K = K;
    A.trees[t].kdmax = kdmax;
    A.trees[t].nleaves = lrank[n];
    A.trees[t].K1 = K1;
  }
}

__global__ __launch_bounds__(kPrepThreads) void cost_prefix_kernel(CostPrefixArgs A) {
  __shared__ int32_t lds_tmp[kPrepThreads];
  const TreeMeta meta = A.trees[A.first_tree + blockIdx.x];
  const int32_t n = meta.n;
  const int32_t* lab = A.label + meta.node_off;
  // A pool tree added before the cost model may hold labels outside its
  // alphabet; such a tree's sums are meaningless (pairs using it are
  // rejected on the host) but its reads stay inside the cost arrays.
  const uint32_t amax = (uint32_t)A.alphabet - 1;
  for (int pass = 0; pass < 2; ++pass) {
    const int32_t* cost = pass ? A.ins : A.del;
    int32_t* out = (pass ? A.pins : A.pdel) + meta.node_off;
    // Each thread owns a contiguous chunk of nodes; a scan of the chunk sums.
    const int per = (n + kPrepThreads - 1) / kPrepThreads;
    const int lo = min(n, (int)threadIdx.x * per), hi = min(n, lo + per);
    int32_t s = 0;
    for (int i = lo; i < hi; ++i) s += cost[min((uint32_t)lab[i], amax)];
    lds_tmp[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kPrepThreads; off <<= 1) {
      const int32_t v = (int)threadIdx.x >= off ? lds_tmp[threadIdx.x - off] : 0;
      __syncthreads();
      lds_tmp[threadIdx.x] += v;
      __syncthreads();
    }
    int32_t run = threadIdx.x ? lds_tmp[threadIdx.x - 1] : 0;
    for (int i = lo; i < hi; ++i) {
      out[i] = run;
      run += cost[min((uint32_t)lab[i], amax)];
    }
    if (threadIdx.x == kPrepThreads - 1) {
      out[n] = lds_tmp[kPrepThreads - 1];
      A.totals[2 * blockIdx.x + pass] = lds_tmp[kPrepThreads - 1];
    }
    __syncthreads();
  }
  const int32_t* orl = A.orl + meta.node_off;
  const int32_t* pd = A.pdel + meta.node_off;
  const int32_t* pi = A.pins + meta.node_off;
  for (int i = threadIdx.x; i < n; i += kPrepThreads) {
    const int32_t e = orl[i] + 1;
    A.rec[meta.node_off + i] = CostRec{pd[i], pi[i], pd[e], pi[e]};
  }
}

void launch_cost_prefix(const CostPrefixArgs& a, int ntrees, hipStream_t s) {
  if (ntrees <= 0) return;
  hipLaunchKernelGGL(cost_prefix_kernel, dim3(ntrees), dim3(kPrepThreads), 0, s, a);
}

void launch_tree_prep(const PrepArgs& a, int ntrees, hipStream_t s) {
  if (ntrees <= 0) return;
  hipLaunchKernelGGL(tree_prep_kernel, dim3(ntrees), dim3(kPrepThreads), 0, s, a);
}

}  // namespace xted
