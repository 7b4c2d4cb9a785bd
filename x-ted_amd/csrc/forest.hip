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

#include "kernels.hpp"

namespace xted {

constexpr int kBig = INT_MAX / 4;

// DPP control codes (GFX9 encoding).
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_WAVE_SHR1 0x138
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

// Inclusive prefix minimum over the 64 lanes of a wave (lane order).
__device__ __forceinline__ int wave_incl_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(kBig, v, DPP_ROW_SHR(1), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(kBig, v, DPP_ROW_SHR(2), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(kBig, v, DPP_ROW_SHR(4), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(kBig, v, DPP_ROW_SHR(8), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(kBig, v, DPP_ROW_BCAST15, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(kBig, v, DPP_ROW_BCAST31, 0xc, 0xf, false));
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
__global__ __launch_bounds__(256) void lane_table_kernel(DPArgs A) {
  const int64_t tix = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (tix >= A.ntasks) return;
  const uint64_t task = A.tasks[tix];
  const PairMeta pm = A.pairs[task_pair(task)];
  const TreeMeta tx = A.pool.trees[pm.tx], ty = A.pool.trees[pm.ty];
  const int32_t i0 = (int32_t)task_i0(task), j0 = (int32_t)task_j0(task);
  const int32_t* orlx = A.pool.orl + tx.node_off;
  const int32_t* orly = A.pool.orl + ty.node_off;
  const int32_t a = orlx[i0] - i0 + 1, b = orly[j0] - j0 + 1;
  const bool rows_x = a <= b;
  const int64_t offR = rows_x ? tx.node_off : ty.node_off;
  const int64_t offC = rows_x ? ty.node_off : tx.node_off;
  const int32_t* labR = A.pool.label + offR;
  const int32_t* orlC = A.pool.orl + offC;
  const int32_t* labC = A.pool.label + offC;
  const int32_t r0 = rows_x ? i0 : j0, c0 = rows_x ? j0 : i0;
  // TD in [r][c] and [c][r] layouts for this orientation.
  int32_t* TDrc = A.td + (rows_x ? pm.td_off : pm.tdt_off);
  int32_t* TDcr = A.td + (rows_x ? pm.tdt_off : pm.td_off);
  const int64_t ldrc = rows_x ? ty.n : tx.n;
  const int64_t ldcr = rows_x ? tx.n : ty.n;
  const int32_t cmax = orlC[c0] + 1;
  const int32_t orl_c0 = cmax - 1;
  const int32_t labr = labR[r0];
  int32_t g = 1;  // F[r0][cmax] = imax - r0
  for (int32_t c = cmax - 1; c >= c0; --c) {
    const int32_t t = cmax - 1 - c;
    const int32_t oc = orlC[c];
    int32_t d;
    const bool tree = (oc == orl_c0);
    if (tree) {
      d = t + (labr != labC[c] ? 1 : 0);
    } else {
      d = (cmax - oc - 1) + TDcr[(int64_t)c * ldcr + r0];
    }
    const int32_t av = min(d, t + 2);
    g = min(av, g + 1);
    if (tree) {
      TDrc[(int64_t)r0 * ldrc + c] = g;
      TDcr[(int64_t)c * ldcr + r0] = g;
    }
  }
}

// ---------------------------------------------------------------------------
// Workgroup row scan: one table per workgroup of W waves; lane l of wave w
// owns columns t in [(w*64+l)*C, (w*64+l)*C + C) (t counted from jmax).
// ---------------------------------------------------------------------------
template <int W, int C>
__global__ __launch_bounds__(64 * W) void rowscan_kernel(DPArgs A) {
  extern __shared__ uint16_t slots[];
  __shared__ int32_t xtot[W];  // per-wave scan totals
  __shared__ int32_t xbnd[W];  // per-wave last-column F value of the row

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
  const int32_t orl_r0 = rmax - 1, orl_c0 = cmax - 1;
  const int32_t lr_c0 = lrankC[c0];
  const int32_t width = lrankC[cmax] - lr_c0 - 1;

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int32_t tbase = (wave * 64 + lane) * C;

  // Per-column state, in registers. Column e of this lane is t = tbase + e,
  // c = cmax-1-t; it is valid while t < ncols. fcol = orl_C(c)+1, so the
  // column is a tree column exactly when fcol == cmax. fw packs the compact
  // far index (high 16 bits) and 1 + the saved-row write index (low 16 bits,
  // 0 = not a leaf+1 column). Fp holds F[r+1][c]; F[r+1][c+1] is Fp[e-1],
  // or fpr0 for e = 0.
  int32_t labc[C], fcol[C], fw[C], Fp[C], tdc[C], tdn[C];
#pragma unroll
  for (int e = 0; e < C; ++e) {
    const int32_t t = tbase + e;
    const int32_t c = cmax - 1 - t;
    labc[e] = 0;
    fcol[e] = cmax;
    fw[e] = 0;
    if (t < ncols) {
      const int32_t oc = orlC[c];
      labc[e] = labC[c];
      fcol[e] = oc + 1;
      const int32_t fidx = (oc == orl_c0) ? 0 : lrankC[oc] - lr_c0;
      const int32_t widx1 = (c > c0 && orlC[c - 1] == c - 1) ? lrankC[c - 1] - lr_c0 + 1 : 0;
      fw[e] = (fidx << 16) | widx1;
    }
    Fp[e] = t + 1;  // F[rmax][c] = cmax - c
    tdc[e] = 0;
    tdn[e] = 0;
  }
  int32_t fpr0 = tbase;  // F[rmax][c+1] for e = 0
  {
    const int32_t* row = TDo + (int64_t)(rmax - 1) * ldo;
#pragma unroll
    for (int e = 0; e < C; ++e)
      if (tbase + e < ncols) tdc[e] = row[cmax - 1 - (tbase + e)];
  }

  for (int32_t r = rmax - 1; r >= r0; --r) {
    // Prefetch the next row's TD values.
    if (r > r0) {
      const int32_t* row = TDo + (int64_t)(r - 1) * ldo;
#pragma unroll
      for (int e = 0; e < C; ++e)
        if (tbase + e < ncols) tdn[e] = row[cmax - 1 - (tbase + e)];
    }
    const int32_t orr = orlR[r];
    const int32_t labr = labR[r];
    const bool tree_row = (orr == orl_r0);
    const int32_t Rf = orr + 1;
    const int32_t fslot = (Rf < rmax) ? kdepR[r] - base - 1 : 0;
    const bool wrow = (r > r0) && (orlR[r - 1] == r - 1);
    const int32_t wslot = wrow ? kdepR[r - 1] - base - 1 : 0;
    const uint16_t* frow = slots + fslot * width;

    int32_t h[C];
#pragma unroll
    for (int e = 0; e < C; ++e) {
      const int32_t t = tbase + e;
      const bool tree_col = (fcol[e] == cmax);
      int32_t d;
      if (tree_row && tree_col) {
        const int32_t fpr = e ? Fp[e - 1] : fpr0;
        d = fpr + (labr != labc[e] ? 1 : 0);
      } else {
        int32_t farv;
        if (Rf == rmax) {
          farv = cmax - fcol[e];
        } else if (tree_col) {
          farv = rmax - Rf;
        } else {
          farv = (int32_t)frow[(uint32_t)fw[e] >> 16];
        }
        d = farv + tdc[e];
      }
      const int32_t av = min(d, Fp[e] + 1);
      h[e] = (t < ncols) ? av - t : kBig;
    }
#pragma unroll
    for (int e = 1; e < C; ++e) h[e] = min(h[e], h[e - 1]);
    const int32_t incl = wave_incl_min(h[C - 1]);
    int32_t carry = min((rmax - r) + 1, wave_shr1(incl, kBig));
    if constexpr (W > 1) {
      if (lane == 63) xtot[wave] = incl;
      __syncthreads();
      for (int w = 0; w < wave; ++w) carry = min(carry, xtot[w]);
    }
    // h becomes G = F[r][c].
#pragma unroll
    for (int e = 0; e < C; ++e) h[e] = tbase + e + min(carry, h[e]);

    // Tree cells -> TD / TD^T; leaf+1 rows -> LDS slots.
    uint16_t* wrow_ptr = slots + wslot * width;
#pragma unroll
    for (int e = 0; e < C; ++e) {
      if (tbase + e >= ncols) continue;
      const int32_t c = cmax - 1 - (tbase + e);
      if (tree_row && fcol[e] == cmax) {
        TDo[(int64_t)r * ldo + c] = h[e];
        TDt[(int64_t)c * ldt + r] = h[e];
      }
      const int32_t widx1 = fw[e] & 0xFFFF;
      if (wrow && widx1) wrow_ptr[widx1 - 1] = (uint16_t)h[e];
    }

    int32_t left_in = rmax - r;  // F[r][cmax]
    if constexpr (W > 1) {
      if (lane == 63) xbnd[wave] = h[C - 1];
      __syncthreads();
      if (wave > 0) left_in = xbnd[wave - 1];
    }
    fpr0 = wave_shr1(h[C - 1], left_in);
#pragma unroll
    for (int e = 0; e < C; ++e) {
      Fp[e] = h[e];
      tdc[e] = tdn[e];
    }
  }
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
  if (cls == kClassLane) {
    const unsigned blocks = (unsigned)((a.ntasks + 255) / 256);
    hipLaunchKernelGGL(lane_table_kernel, dim3(blocks), dim3(256), 0, s, a);
    return true;
  }
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
