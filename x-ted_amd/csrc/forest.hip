// Forest-distance DP kernels for gfx950 (wave64).
//
// One keyroot-pair table (i0, j0) fills F over rows i in [i0, imax] and columns
// j in [j0, jmax], imax = orl_x(i0)+1, jmax = orl_y(j0)+1, with
//   F[imax][j] = jmax - j,  F[i][jmax] = imax - i,
//   tree cell (orl_x(i)==orl_x(i0) && orl_y(j)==orl_y(j0)):
//     F[i][j] = min(F[i+1][j+1] + (lab_x(i)!=lab_y(j)), F[i+1][j]+1, F[i][j+1]+1) = TD[i][j]
//   forest cell:
//     F[i][j] = min(F[orl_x(i)+1][orl_y(j)+1] + TD[i][j], F[i+1][j]+1, F[i][j+1]+1)
// (X-TED_CPU/parallel_computing.cpp:4-45, X-TED_GPU/parallel_X_TED.cu:5-55,97-122).
//
// The tables are computed row by row along the table's short side ("rows"),
// each row as one min-plus prefix scan along the long side ("columns"):
//   with t = jmax-1-j and a[t] = min(diag-or-far term, F[i+1][j]+1),
//   F[i][j] = t + min(imax-i+1, min_{t'<=t} (a[t'] - t')).
// Integer min/plus are associative, so this re-association is bit-exact.
// The transposed problem (rows = y) is the same recurrence on TD^T; both TD
// and TD^T are kept so each orientation reads its rows contiguously.
//
// Far reads F[orl_x(i)+1][orl_y(j)+1] only touch rows r with r-1 a leaf and
// columns c with c-1 a leaf. Such a row is written once and read by the rows
// of one right path, i.e. while the sweep is inside one keyroot's subtree, so
// it lives in an LDS slot indexed by that keyroot's nesting depth below the
// table's keyroot; columns are compacted by leaf rank. Values are < 2^16 and
// stored as uint16.
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "kernels.hpp"

namespace xted {

constexpr int kBig = INT_MAX / 4;

// DPP control codes (GFX9 encoding).
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_WAVE_SHR1 0x138
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

// Inclusive prefix minimum over the 64 lanes of a wave (lane order). A lane
// whose DPP source is out of range keeps its own value as the "old"
// operand, which is the identity of min(v, v).
__device__ __forceinline__ int wave_incl_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(v, v, DPP_ROW_SHR(1), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, DPP_ROW_SHR(2), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, DPP_ROW_SHR(4), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, DPP_ROW_SHR(8), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, DPP_ROW_BCAST15, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, DPP_ROW_BCAST31, 0xc, 0xf, false));
  return v;
}

// Value of lane-1 (lane 0 receives `fill`).
__device__ __forceinline__ int wave_shr1(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, DPP_WAVE_SHR1, 0xf, 0xf, false);
}

// ---------------------------------------------------------------------------
// Lane-per-table: tables whose short side is a single node. The row is the
// single leaf r0 (so every far row is the boundary row), the lane walks the
// columns of the other subtree right to left carrying F[r0][c+1].
// Tasks with the same column keyroot sit next to each other, so the orl and
// label loads of one wave are mostly uniform.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lane_table(const PoolView& P, const PairMeta& pm, int32_t i0,
                                           int32_t j0, int32_t* td) {
  const TreeMeta tx = P.trees[pm.tx], ty = P.trees[pm.ty];
  const int32_t* orlx = P.orl + tx.node_off;
  const int32_t* orly = P.orl + ty.node_off;
  const int32_t a = orlx[i0] - i0 + 1, b = orly[j0] - j0 + 1;
  const bool rows_x = a <= b;
  const int64_t offR = rows_x ? tx.node_off : ty.node_off;
  const int64_t offC = rows_x ? ty.node_off : tx.node_off;
  const int32_t* labR = P.label + offR;
  const int32_t* orlC = P.orl + offC;
  const int32_t* labC = P.label + offC;
  const int32_t r0 = rows_x ? i0 : j0, c0 = rows_x ? j0 : i0;
  // TD in [r][c] and [c][r] layouts for this orientation.
  int32_t* TDrc = td + (rows_x ? pm.td_off : pm.tdt_off);
  int32_t* TDcr = td + (rows_x ? pm.tdt_off : pm.td_off);
  const int64_t ldrc = rows_x ? ty.n : tx.n;
  const int64_t ldcr = rows_x ? tx.n : ty.n;
  const int32_t cmax = orlC[c0] + 1;
  const int32_t orl_c0 = cmax - 1;
  const int32_t labr = labR[r0];
  // Loads run kAhead columns ahead of the recurrence: ring slot k holds
  // column c-k's orl, label and TD value. The main loop (c - kAhead >= c0)
  // walks pointers and never clamps; the last kAhead columns are already in
  // the ring. Values are only combined when used, so no load waits inside a
  // branch, and the main loop is unrolled by the ring depth so the rotation
  // is register renaming.
  constexpr int kAhead = 4;
  int32_t ring_o[kAhead], ring_l[kAhead], ring_v[kAhead];
#pragma unroll
  for (int k = 0; k < kAhead; ++k) {
    const int32_t c = max(cmax - 1 - k, c0);
    ring_o[k] = orlC[c];
    ring_l[k] = labC[c];
    ring_v[k] = TDcr[(int64_t)c * ldcr + r0];
  }
  const int32_t* po = orlC + (cmax - 1 - kAhead);
  const int32_t* pl = labC + (cmax - 1 - kAhead);
  const int32_t* pv = TDcr + (int64_t)(cmax - 1 - kAhead) * ldcr + r0;
  int32_t* pw_cr = TDcr + (int64_t)(cmax - 1) * ldcr + r0;
  int32_t* pw_rc = TDrc + (int64_t)r0 * ldrc + (cmax - 1);
  int32_t g = 1;  // F[r0][cmax] = imax - r0
  int32_t t = 0;
  auto step = [&]() {
    const int32_t oc = ring_o[0], lab = ring_l[0], v = ring_v[0];
#pragma unroll
    for (int k = 0; k + 1 < kAhead; ++k) {
      ring_o[k] = ring_o[k + 1];
      ring_l[k] = ring_l[k + 1];
      ring_v[k] = ring_v[k + 1];
    }
    const bool tree = (oc == orl_c0);
    // tree cell: F[imax][c+1] + rename = t + (lab differ); forest cell:
    // F[imax][orl(c)+1] + TD[r0][c] = (cmax - orl(c) - 1) + TD. Both terms
    // are formed and one is selected, so the step has no divergent branch.
    const int32_t dt = t + (int32_t)(labr != lab);
    const int32_t df = (cmax - oc - 1) + v;
    const int32_t d = tree ? dt : df;
    g = min(min(d, t + 2), g + 1);
    if (tree) {
      *pw_rc = g;
      *pw_cr = g;
    }
    pw_rc -= 1;
    pw_cr -= ldcr;
    ++t;
  };
  const int32_t ncols = cmax - c0;
#pragma unroll 4
  for (int32_t i = 0; i < ncols - kAhead; ++i) {
    const int32_t o = *po, l = *pl, v = *pv;
    step();
    ring_o[kAhead - 1] = o;
    ring_l[kAhead - 1] = l;
    ring_v[kAhead - 1] = v;
    po -= 1;
    pl -= 1;
    pv -= ldcr;
  }
#pragma unroll
  for (int k = 0; k < kAhead; ++k)
    if (t < ncols) step();
}

// Single-row tables of one level, without a task list: thread g finds its
// pair in the level's prefix over pairs, then its cell of rectangle A
// (consecutive lanes share the y keyroot) or B (consecutive lanes share the
// x keyroot).
__global__ __launch_bounds__(256) void lane_level_kernel(LaneArgs A, unsigned long long total) {
  const unsigned long long g = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const unsigned long long* off = A.offsets + (int64_t)A.level * (A.npairs + 1);
  int lo = 0, hi = A.npairs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const int p = lo;
  int32_t sz[5];
  (void)lane_rect_count(A, p, A.level, sz);
  const PairMeta pm = A.pairs[p];
  const TreeMeta tx = A.pool.trees[pm.tx], ty = A.pool.trees[pm.ty];
  int32_t q = (int32_t)(g - off[p]);
  int32_t k, l;
  if (q < sz[0]) {
    k = q % tx.K1;
    l = sz[3] + q / tx.K1;
  } else {
    q -= sz[0];
    l = q % ty.K1;
    k = sz[4] + q / ty.K1;
  }
  lane_table(A.pool, pm, A.pool.skr[tx.node_off + k], A.pool.skr[ty.node_off + l], A.td);
}

void launch_lane_level(const LaneArgs& a, unsigned long long total, hipStream_t s) {
  if (total == 0) return;
  hipLaunchKernelGGL(lane_level_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a,
                     total);
}

// ---------------------------------------------------------------------------
// Workgroup row scan: one table per workgroup of W waves, C columns per
// lane. Columns are counted from jmax: t = jmax-1-j. Wave w owns the block
// t in [w*64*C, (w+1)*64*C). Inside it, lanes own E contiguous columns per
// group and the G = C/E groups are interleaved: lane l's element (g, e) is
//   t = w*64*C + g*64*E + l*E + e,   E = min(C, 4).
// So one group of a TD row is a 64*E-int contiguous span that the wave
// reads with one coalesced E-int load per lane, and the row scan is G
// chained wave scans over groups.
// ---------------------------------------------------------------------------
typedef int32_t int4u __attribute__((ext_vector_type(4), aligned(4)));
typedef int32_t int2u __attribute__((ext_vector_type(2), aligned(4)));

template <int W, int C>
__global__ __launch_bounds__(64 * W) void rowscan_kernel(DPArgs A) {
  constexpr int E = C <= 4 ? C : 4;  // contiguous columns per group
  constexpr int G = C / E;           // groups per lane
  // TD ring slots (a power of two dividing 64): up to 32 registers of ring,
  // so narrow lanes prefetch more rows ahead.
  constexpr int S = C >= 16 ? 2 : C >= 8 ? 4 : C >= 4 ? 8 : 16;
  constexpr int PD = S - 1;  // rows of TD prefetched
  extern __shared__ uint16_t slots[];
  // Per-wave scan totals, double-buffered by row parity: a wave may write
  // row r-1's total while a slower wave still reads row r's.
  __shared__ int32_t xtot[2][W];

  const int64_t tix = blockIdx.x;
  if (tix >= A.ntasks) return;
  const uint64_t task = A.tasks[tix];
  const PairMeta pm = A.pairs[task_pair(task)];
  const TreeMeta tx = A.pool.trees[pm.tx], ty = A.pool.trees[pm.ty];
  const int32_t i0 = (int32_t)task_i0(task), j0 = (int32_t)task_j0(task);
  const int32_t a = A.pool.orl[tx.node_off + i0] - i0 + 1;
  const int32_t b = A.pool.orl[ty.node_off + j0] - j0 + 1;
  const bool rows_x = a <= b;
  const int64_t offR = rows_x ? tx.node_off : ty.node_off;
  const int64_t offC = rows_x ? ty.node_off : tx.node_off;
  const int32_t* __restrict__ orlR = A.pool.orl + offR;
  const int32_t* __restrict__ labR = A.pool.label + offR;
  const int32_t* __restrict__ kdepR = A.pool.kdep + offR;
  const int32_t* __restrict__ orlC = A.pool.orl + offC;
  const int32_t* __restrict__ labC = A.pool.label + offC;
  const int32_t* __restrict__ lrankC = A.pool.lrank + offC;
  const int32_t r0 = rows_x ? i0 : j0, c0 = rows_x ? j0 : i0;
  int32_t* TDo = A.td + (rows_x ? pm.td_off : pm.tdt_off);  // [r][c]
  int32_t* TDt = A.td + (rows_x ? pm.tdt_off : pm.td_off);  // [c][r]
  const int64_t ldo = rows_x ? ty.n : tx.n;
  const int64_t ldt = rows_x ? tx.n : ty.n;

  const int32_t rmax = orlR[r0] + 1, cmax = orlC[c0] + 1;
  const int32_t ncols = cmax - c0;
  const int32_t base = kdepR[r0];
  const int32_t orl_c0 = cmax - 1;
  const int32_t lr_c0 = lrankC[c0];
  // A slot row holds the compact far columns (leaves of subtree(c0) but the
  // last) and, at index width, F[R][cmax] = rmax - R for the tree columns.
  const int32_t width = lrankC[cmax] - lr_c0 - 1;
  const int32_t rowlen = width + 1;

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int32_t tlane = wave * 64 * C + lane * E;  // t of this lane's element (0, 0)
#define TCOL(g, e) (tlane + (g) * 64 * E + (e))

  // Per-column state, element j = g*E + e (t = TCOL(g, e), c = cmax-1-t).
  //   bfar: F[rmax][orl_C(c)+1] = cmax - orl_C(c) - 1 (the far value on the
  //         boundary row), or -1 on a tree column (orl_C(c) == orl_C(c0));
  //   labc: the column label (tree cells only);
  //   foff: the column's far index in a slot row (width for tree columns);
  //   woff: the index this column writes in a saved row (c-1 a leaf of
  //         subtree(c0)), or -1;
  //   Fpm:  F[r+1][c] + 1 - t (the scan works on a[t] - t);
  //   negt: -t.
  // Columns past the table (t >= ncols) come after every real column in the
  // scan order, so whatever they hold never reaches a real column; only
  // their stores are masked.
  int32_t bfar[C], labc[C], foff[C], woff[C], Fpm[C], negt[C], tdq[S][C];
  int32_t fpr0[G];  // F[r+1][c+1] of each group's element e = 0
#pragma unroll
  for (int g = 0; g < G; ++g) {
    fpr0[g] = TCOL(g, 0);  // F[rmax][c+1] = cmax-1-c = t
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = g * E + e;
      const int32_t t = TCOL(g, e);
      const int32_t c = max(cmax - 1 - t, c0);
      const int32_t oc = orlC[c];
      const int32_t lab = labC[c];
      const int32_t lro = lrankC[oc];
      const int32_t ocm = orlC[max(c - 1, c0)];
      const int32_t lrm = lrankC[max(c - 1, c0)];
      const bool tree = (oc == orl_c0);
      bfar[j] = tree ? -1 : cmax - oc - 1;
      labc[j] = lab;
      foff[j] = tree ? width : lro - lr_c0;
      woff[j] = (t < ncols && c > c0 && ocm == c - 1) ? lrm - lr_c0 : -1;
      Fpm[j] = 2;  // F[rmax][c] + 1 - t = cmax - c + 1 - t
      negt[j] = -t;
    }
  }

  // TD row rr of this lane's columns into out[]. Loads are unconditional, so
  // the compiler never waits for them inside a branch: a group with no valid
  // column reads the row's first E ints instead, a partly valid group reads
  // up to E-1 ints before column c0 (the arena keeps kTdPad ints of padding
  // before and after every table, so both stay inside the allocation).
  auto load_row = [&](int32_t rr, int32_t* out) {
    const int32_t* row = TDo + (int64_t)rr * ldo;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int32_t tb = TCOL(g, 0);
      const int32_t off = tb < ncols ? cmax - E - tb : c0;
      if constexpr (E == 4) {
        const int4u v = *reinterpret_cast<const int4u*>(row + off);
        out[g * 4 + 0] = v.w;
        out[g * 4 + 1] = v.z;
        out[g * 4 + 2] = v.y;
        out[g * 4 + 3] = v.x;
      } else if constexpr (E == 2) {
        const int2u v = *reinterpret_cast<const int2u*>(row + off);
        out[g * 2 + 0] = v.y;
        out[g * 2 + 1] = v.x;
      } else {
        out[g] = row[off];
      }
    }
  };
#pragma unroll
  for (int q = 0; q < PD; ++q) load_row(max(rmax - 1 - q, r0), tdq[q]);

  // Row metadata, 64 rows per chunk: lane i holds row rmax-1-(64q+i) as
  // (label, Rf = orl_R(r)+1, slots = fslot | wslot << 16 | wrow << 31), read
  // per row with readlane; the next chunk is loaded one chunk ahead.
  auto load_meta = [&](int q, int32_t& mlab, int32_t& mrf, int32_t& mslot) {
    const int32_t rr = rmax - 1 - (64 * q + lane);
    const int32_t r = max(rr, r0);  // rows past the table read row r0
    const int32_t rp = max(r - 1, r0);
    const int32_t orr = orlR[r], lab = labR[r], kd = kdepR[r];
    const int32_t orp = orlR[rp], kdp = kdepR[rp];
    const int32_t fs = (orr + 1 < rmax) ? kd - base - 1 : 0;
    const int32_t ws =
        (r > r0 && orp == r - 1) ? (((kdp - base - 1) << 16) | (int32_t)0x80000000) : 0;
    mlab = lab;
    mrf = orr + 1;
    mslot = fs | ws;
  };
  const int32_t nrows = rmax - r0;
  const int32_t nchunks = (nrows + 63) >> 6;
  int32_t mlab, mrf, mslot, nlab = 0, nrf = 0, nslot = 0;
  load_meta(0, mlab, mrf, mslot);
  if (nchunks > 1) load_meta(1, nlab, nrf, nslot);

  for (int32_t q = 0; q < nchunks; ++q) {
    const int32_t rows_here = min(64, nrows - 64 * q);
    // Rows go S at a time with the inner loop fully unrolled, so row k of the
    // table always uses ring slot k % S and every slot index is static: row
    // r reads slot u and prefetches row r-PD into slot (u+PD) % S, the slot
    // the previous row has just finished with.
    for (int32_t i0r = 0; i0r < rows_here; i0r += S) {
#pragma unroll
      for (int u = 0; u < S; ++u) {
        const int32_t i = i0r + u;
        if (i >= rows_here) break;
        const int32_t r = rmax - 1 - (64 * q + i);
        load_row(max(r - PD, r0), tdq[(u + PD) % S]);

        const int32_t labr = __builtin_amdgcn_readlane(mlab, i);
        const int32_t Rf = __builtin_amdgcn_readlane(mrf, i);
        const int32_t sl = __builtin_amdgcn_readlane(mslot, i);
        const bool tree_row = (Rf == rmax);
        const uint16_t* frow = slots + (sl & 0xFFFF) * rowlen;
        const bool wrow = sl < 0;
        uint16_t* wrow_ptr = slots + ((sl >> 16) & 0x7FFF) * rowlen;

        // h = a[t] - t per element, local prefix per group, chained wave
        // scans. Tree rows (orl_R(r) == orl_R(r0)) and the other rows take
        // separate code paths on the uniform flag.
        int32_t h[C], pre[G];
        int32_t carry_w = kBig;  // min over this wave's earlier groups
        auto row_terms = [&](auto tree_tag) {
          constexpr bool kTree = decltype(tree_tag)::value;
          int32_t sv[C];
          if constexpr (!kTree) {
#pragma unroll
            for (int j = 0; j < C; ++j) sv[j] = frow[foff[j]];
          }
#pragma unroll
          for (int g = 0; g < G; ++g) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
              const int j = g * E + e;
              int32_t dm;  // diagonal-or-far term minus t
              if constexpr (kTree) {
                // diagonal F[r+1][c+1] + rename on a tree column, else the
                // boundary-row far value. F[r+1][c+1] - t = Fpm[j-1] - 2.
                const int32_t fprm = e ? Fpm[j - 1] - 2 : fpr0[g] + negt[j];
                dm = bfar[j] < 0 ? fprm + (int32_t)(labr != labc[j]) : bfar[j] + tdq[u][j] + negt[j];
              } else {
                dm = sv[j] + tdq[u][j] + negt[j];
              }
              h[j] = min(dm, Fpm[j]);
              if (e) h[j] = min(h[j], h[j - 1]);
            }
            const int32_t incl = wave_incl_min(h[g * E + E - 1]);
            pre[g] = min(carry_w, wave_shr1(incl, kBig));
            carry_w = min(carry_w, __builtin_amdgcn_readlane(incl, 63));
          }
        };
        if (tree_row) row_terms(std::true_type{});
        else row_terms(std::false_type{});

        int32_t carry = (rmax - r) + 1;  // F[r][cmax] - (-1)
        if constexpr (W > 1) {
          int32_t* xt = xtot[r & 1];
          if (lane == 0) xt[wave] = carry_w;
          __syncthreads();
          for (int w = 0; w < wave; ++w) carry = min(carry, xt[w]);
        }
        // h becomes G - t; Fpm = G + 1 - t = h + 1.
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int32_t cg = min(carry, pre[g]);
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int j = g * E + e;
            h[j] = min(cg, h[j]);
            Fpm[j] = h[j] + 1;
          }
        }

        // Tree cells -> TD / TD^T (tree rows only); saved rows -> LDS slots.
        if (tree_row) {
#pragma unroll
          for (int j = 0; j < C; ++j) {
            const int32_t t = TCOL(j / E, j % E);
            if (t < ncols && bfar[j] < 0) {
              const int32_t c = cmax - 1 - t;
              TDo[(int64_t)r * ldo + c] = h[j] + t;
              TDt[(int64_t)c * ldt + r] = h[j] + t;
            }
          }
        }
        if (wrow) {
#pragma unroll
          for (int j = 0; j < C; ++j)
            if (woff[j] >= 0) wrow_ptr[woff[j]] = (uint16_t)(h[j] - negt[j]);
          if (threadIdx.x == 0) wrow_ptr[width] = (uint16_t)(rmax - r);
          // The next row may read this slot row at columns of other waves.
          // (Every other LDS reuse is ordered by the barrier above: a slot
          // row is only rewritten after the rows that read it.)
          if constexpr (W > 1) __syncthreads();
        }

        // F[r][c] at the column just before this wave's block (t = w*64*C-1).
        // The previous wave's last element ends with h = min(its carry-in,
        // its own total) = this wave's carry-in, and F = h + t; for wave 0
        // this gives F[r][cmax] = rmax - r.
        const int32_t left_in = carry + (wave * 64 * C - 1);
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int jl = g * E + E - 1;  // the group's last column
          const int32_t fill =
              g ? __builtin_amdgcn_readlane(h[jl - E] - negt[jl - E], 63) : left_in;
          fpr0[g] = wave_shr1(h[jl] - negt[jl], fill);
        }
      }
    }
    mlab = nlab;
    mrf = nrf;
    mslot = nslot;
    if (q + 2 < nchunks) load_meta(q + 2, nlab, nrf, nslot);
  }
#undef TCOL
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
typedef void (*RowKernel)(DPArgs);

template <int W, int C>
static void launch_row(const DPArgs& a, int lds_bytes, hipStream_t s) {
  hipLaunchKernelGGL((rowscan_kernel<W, C>), dim3((unsigned)a.ntasks), dim3(64 * W),
                     (size_t)lds_bytes, s, a);
}

template <int W>
static void launch_row_c(int ci, const DPArgs& a, int lds, hipStream_t s) {
  switch (ci) {
    case 0: launch_row<W, 1>(a, lds, s); break;
    case 1: launch_row<W, 2>(a, lds, s); break;
    case 2: launch_row<W, 4>(a, lds, s); break;
    case 3: launch_row<W, 8>(a, lds, s); break;
    default: launch_row<W, 16>(a, lds, s); break;
  }
}

// Grants every row-scan instantiation the LDS it may ask for (> 64 KiB needs
// an explicit attribute).
template <int W, int C>
static hipError_t allow_lds(int limit) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&rowscan_kernel<W, C>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, limit);
}
template <int W>
static hipError_t allow_lds_c(int limit) {
  hipError_t e;
  if ((e = allow_lds<W, 1>(limit)) != hipSuccess) return e;
  if ((e = allow_lds<W, 2>(limit)) != hipSuccess) return e;
  if ((e = allow_lds<W, 4>(limit)) != hipSuccess) return e;
  if ((e = allow_lds<W, 8>(limit)) != hipSuccess) return e;
  return allow_lds<W, 16>(limit);
}
hipError_t forest_init_attributes(int limit) {
  hipError_t e;
  if ((e = allow_lds_c<1>(limit)) != hipSuccess) return e;
  if ((e = allow_lds_c<2>(limit)) != hipSuccess) return e;
  if ((e = allow_lds_c<4>(limit)) != hipSuccess) return e;
  if ((e = allow_lds_c<8>(limit)) != hipSuccess) return e;
  return allow_lds_c<16>(limit);
}

// Launches one (class, task list). Returns false for a class without a kernel.
bool launch_forest(int cls, const DPArgs& a, int lds_bytes, hipStream_t s) {
  if (a.ntasks <= 0) return true;
  if (cls >= 1 && cls <= kNumRowClasses) {
    const int wi = (cls - 1) / 5, ci = (cls - 1) % 5;
    switch (wi) {
      case 0: launch_row_c<1>(ci, a, lds_bytes, s); break;
      case 1: launch_row_c<2>(ci, a, lds_bytes, s); break;
      case 2: launch_row_c<4>(ci, a, lds_bytes, s); break;
      case 3: launch_row_c<8>(ci, a, lds_bytes, s); break;
      default: launch_row_c<16>(ci, a, lds_bytes, s); break;
    }
    return true;
  }
  return false;
}

}  // namespace xted
