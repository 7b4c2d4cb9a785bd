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

#include <algorithm>
#include <climits>
#include <type_traits>
#include <utility>

#include "kernels.hpp"

namespace xted {

constexpr int kBig = INT_MAX / 4;

// DPP control codes (GFX9 encoding).
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_WAVE_SHR1 0x138
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

// Inclusive prefix minimum over the 64 lanes of a wave (lane order). A lane
// whose DPP source is out of range (or whose row the mask leaves out) reads
// INT_MAX, the identity of min; with that identity as the "old" operand the
// compiler fuses each move into its min (one v_min_i32 with a DPP source).
__device__ __forceinline__ int wave_incl_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_SHR(1), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_SHR(2), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_SHR(4), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_SHR(8), 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_BCAST15, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_BCAST31, 0xc, 0xf, false));
  return v;
}

// Value of lane-1 (lane 0 receives `fill`).
__device__ __forceinline__ int wave_shr1(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, DPP_WAVE_SHR1, 0xf, 0xf, false);
}

// ---------------------------------------------------------------------------
// Leaf-keyroot lines in closed form. A line is a leaf keyroot v of one tree R
// of a pair; over the other tree T it is TD(v, c) = (orl(c) - c + 1) -
// [nxt(c) <= orl(c)], nxt(c) being the first position >= c whose label is
// label(v). The line is a row of TD (v of x) or of TD^T (v of y) and a column
// of the other layout.
//
// One workgroup of 4 waves takes the leaf keyroots among 64 consecutive
// nodes of one side (a tile) and walks T right to left, 64 positions per
// step, lane l at position c = top - l. For each of its lines a wave forms
// nxt as a prefix minimum over lanes plus the minimum carried from the
// positions above, writes the 64 values as one contiguous row segment and
// keeps them in an LDS tile; then the workgroup writes the tile transposed,
// one position per wave instruction with the tile's 64 nodes across the
// lanes: whole 256-byte runs of the column layout. The entries of nodes that
// are not leaf keyroots carry no value yet; they belong to tables of two
// non-leaf keyroots, which run after this kernel and overwrite them, and no
// table reads a cell before the level that writes it. Positions that are
// leaf keyroots of T are left out: their whole row is a line of the other
// side, written by its tiles.
// ---------------------------------------------------------------------------
constexpr int kLeafTile = 64;
__global__ __launch_bounds__(256) void leaf_kernel(LeafArgs A, int64_t ntiles, int64_t base) {
  // Row kLeafTile of the tile takes the values of lines past the tile's
  // count, which are computed (so a wave's lines have no branches between
  // them and their scans interleave) but never stored.
  __shared__ int32_t tile[kLeafTile + 1][kLeafTile + 1];
  __shared__ int32_t lines[kLeafTile];
  const int64_t tix = base + blockIdx.x;
  if (tix >= ntiles) return;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = A.npairs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (A.tile_start[mid] <= tix) lo = mid;
    else hi = mid - 1;
  }
  const PairMeta pm = A.pairs[lo];
  const TreeMeta tx = A.pool.trees[pm.tx], ty = A.pool.trees[pm.ty];
  const int32_t q = (int32_t)(tix - A.tile_start[lo]);
  const int32_t qx = (tx.n + kLeafTile - 1) / kLeafTile;
  const bool of_x = q < qx;  // nodes of x: the lines are rows of TD
  const TreeMeta& R = of_x ? tx : ty;
  const TreeMeta& T = of_x ? ty : tx;
  const int32_t v0 = (of_x ? q : q - qx) * kLeafTile;
  // This lane's node of the tile, and whether it is a leaf keyroot.
  const int32_t vl = v0 + lane;
  const int32_t vc = min(vl, R.n - 1);
  const bool is_line = vl < R.n && A.pool.orl[R.node_off + vc] == vc &&
                       A.pool.kdn[R.node_off + vc] >= 0;
  const unsigned long long lmask = __ballot(is_line);
  const int nl = __popcll(lmask);
  if (nl == 0) return;  // uniform over the workgroup
  if (wave == 0 && is_line) lines[__popcll(lmask & ((1ull << lane) - 1))] = lane;
  const int32_t laml = A.pool.label[R.node_off + vc];
  const int32_t* __restrict__ labT = A.pool.label + T.node_off;
  const int32_t* __restrict__ orlT = A.pool.orl + T.node_off;
  const int32_t* __restrict__ kdnT = A.pool.kdn + T.node_off;
  // of_x: TD[v][c] = td[td_off + v*ny + c], TD^T[c][v] = td[tdt_off + c*nx + v];
  // else: TD^T[v][c] = td[tdt_off + v*nx + c], TD[c][v] = td[td_off + c*ny + v].
  td_t* const rows = A.td + (of_x ? pm.td_off : pm.tdt_off);  // [v][c], stride T.n
  td_t* const cols = A.td + (of_x ? pm.tdt_off : pm.td_off);  // [c][v], stride R.n
  const bool col_ok = vl < R.n;
  __syncthreads();
  constexpr int kPer = kLeafTile / 4;  // lines (and positions) per wave
  int32_t carry[kPer];
#pragma unroll
  for (int m = 0; m < kPer; ++m) carry[m] = INT_MAX;
  // Positions of the next step are read one step ahead.
  int32_t top = T.n - 1;
  int32_t cc = max(top - lane, 0);
  int32_t nlab = labT[cc], norl = orlT[cc], nkdn = kdnT[cc];
  for (; top >= 0; top -= 64) {
    const int32_t c = top - lane;
    const int32_t lab = nlab, oc = norl;
    const bool c_line = oc == c && nkdn >= 0;  // c is a leaf keyroot of T
    if (top >= 64) {
      cc = max(top - 64 - lane, 0);
      nlab = labT[cc];
      norl = orlT[cc];
      nkdn = kdnT[cc];
    }
    const int32_t size = oc - c + 1;
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
      const int k = wave + 4 * m;  // lines interleaved over the waves
      const bool real = k < nl;    // uniform
      const int32_t off = real ? lines[k] : kLeafTile;
      const int32_t lam = __builtin_amdgcn_readlane(laml, real ? off : 0);
      const int32_t incl = wave_incl_min(c >= 0 && lab == lam ? c : INT_MAX);
      const int32_t nxt = min(carry[m], incl);
      carry[m] = min(carry[m], __builtin_amdgcn_readlane(incl, 63));
      const int32_t d = size - (int32_t)(nxt <= oc);
      if (real && c >= 0) rows[(int64_t)(v0 + off) * T.n + c] = (td_t)d;
      tile[off][lane] = d;
    }
    const unsigned long long skip = __ballot(c < 0 || c_line);
    __syncthreads();
    // Transposed: position p = wave + 4m of this step, node v0 + lane.
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
      const int p = wave + 4 * m;
      if (!((skip >> p) & 1ull) && col_ok)
        cols[(int64_t)(top - p) * R.n + vl] = (td_t)tile[lane][p];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Leaf-keyroot lines under a cost model (the closed form in kernels.hpp). For
// line v (a leaf of R) over the other tree T: with H the cost of T's nodes
// that are not mapped (inserted when v is x's, deleted when it is y's), w(k)
// = ren(v, k) - H(k) (ren(k, v) for v of y) and d(v) the cost of v alone,
//   TD(v, c) = H(sub c) + min(d(v), m(c)),   m(c) = min over k in sub(c) of w(k).
// One right-to-left sweep over T gives every m(c): with R(c) the minimum over
// c's subtree and the subtrees of its later siblings,
//   m(c) = w(c) for a leaf, min(w(c), R(c+1)) otherwise;
//   R(c) = min(m(c), R(next sibling)) when c has one, m(c) otherwise.
// Between a node and its previous sibling the sweep passes only the
// previous sibling's subtree, so R of a next sibling waits on a stack: pushed
// at a node that has a previous sibling, popped at a node that has a next
// one. All lines of a tile sweep the same T, so every push and pop is the
// same for all lanes and lane l keeps its line's entries in column l of the
// LDS stack, the top entry in a register (so a push followed by a pop does
// not wait on LDS). One wave per tile of 64 consecutive nodes of R (the unit
// kernel's tiles); lane l is node v0 + l. Per step of 64 positions the
// metadata of position top - i sits in lane i (read per position with
// readlane), the rename values are loaded a chunk of 16 positions ahead,
// and the 64 x 64 values (below 2^16) are written from a 16-bit LDS tile as
// the unit kernel writes them: row segments of every line, then the column
// layout.
// ---------------------------------------------------------------------------
constexpr int kCostLeafChunk = 16;
__global__ __launch_bounds__(64) void leaf_cost_kernel(LeafArgs A, int64_t ntiles, int64_t base) {
  __shared__ int16_t stk[kCostLeafStack][64];
  __shared__ uint16_t tile[kLeafTile][kLeafTile + 2];
  const int64_t tix = base + blockIdx.x;
  if (tix >= ntiles) return;
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = A.npairs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (A.tile_start[mid] <= tix) lo = mid;
    else hi = mid - 1;
  }
  const PairMeta pm = A.pairs[lo];
  const TreeMeta tx = A.pool.trees[pm.tx], ty = A.pool.trees[pm.ty];
  const int32_t q = (int32_t)(tix - A.tile_start[lo]);
  const int32_t qx = (tx.n + kLeafTile - 1) / kLeafTile;
  const bool of_x = q < qx;  // nodes of x: the lines are rows of TD
  const TreeMeta& R = of_x ? tx : ty;
  const TreeMeta& T = of_x ? ty : tx;
  const int32_t v0 = (of_x ? q : q - qx) * kLeafTile;
  const int32_t vl = v0 + lane;
  const int32_t vc = min(vl, R.n - 1);
  const bool is_line = vl < R.n && A.pool.orl[R.node_off + vc] == vc &&
                       A.pool.kdn[R.node_off + vc] >= 0;
  const unsigned long long lmask = __ballot(is_line);
  if (lmask == 0) return;  // uniform (one wave)
  const int32_t labv = A.pool.label[R.node_off + vc];
  // d(v): deleting v (a node of x) or inserting it (a node of y).
  const int32_t* Pv = (of_x ? A.pool.pdel : A.pool.pins) + R.node_off;
  const int32_t dv = Pv[vc + 1] - Pv[vc];
  // T's unmapped nodes are inserted (v of x) or deleted (v of y).
  const int32_t* PT = (of_x ? A.pool.pins : A.pool.pdel) + T.node_off;
  const int32_t* HT = of_x ? A.ins : A.del;
  const int32_t* __restrict__ labT = A.pool.label + T.node_off;
  const int32_t* __restrict__ orlT = A.pool.orl + T.node_off;
  const int32_t* __restrict__ parT = A.pool.parent + T.node_off;
  const int32_t* __restrict__ kdnT = A.pool.kdn + T.node_off;
  // ren(v, k) = ren[labv * a + lab k] for v of x, ren[lab k * a + labv] for v of y.
  const int32_t* const renv = A.ren + (of_x ? (int64_t)labv * A.alphabet : labv);
  const int32_t rstride = of_x ? 1 : A.alphabet;
  td_t* const rows = A.td + (of_x ? pm.td_off : pm.tdt_off);  // [v][c], stride T.n
  td_t* const cols = A.td + (of_x ? pm.tdt_off : pm.td_off);  // [c][v], stride R.n
  const bool col_ok = vl < R.n;
  // Stack entries [0, sp-1) in LDS, entry sp-1 in tos.
  int32_t sp = 0, rprev = 0, tos = 0;
  for (int32_t top = T.n - 1; top >= 0; top -= kLeafTile) {
    // Lane i: position p = top - i (positions past 0 repeat position 0).
    const int32_t p = top - lane;
    const int32_t pc = max(p, 0);
    const int32_t lab = labT[pc], o = orlT[pc], par = parT[pc];
    const int32_t opar = par >= 0 ? orlT[par] : o;
    // bit 0: internal node; bit 1: has a previous sibling (push); bit 2: has
    // a next sibling (pop); bit 3: a leaf keyroot of T (its cells are written
    // by T's own tiles in the column layout)
    const int32_t flags = (o > p ? 1 : 0) | (p > 0 && par != p - 1 ? 2 : 0) |
                          (p > 0 && o < opar ? 4 : 0) | (o == p && kdnT[pc] >= 0 ? 8 : 0);
    const int32_t hsub = PT[o + 1] - PT[pc];  // H(sub p)
    const int32_t hk = HT[lab];               // H(p)
    const int32_t cnt = min(kLeafTile, top + 1);
    int32_t wq[2][kCostLeafChunk];
    auto issue = [&](int chunk, int32_t* dst) {
#pragma unroll
      for (int j = 0; j < kCostLeafChunk; ++j)
        dst[j] = renv[(int64_t)__builtin_amdgcn_readlane(lab, chunk * kCostLeafChunk + j) * rstride];
    };
    issue(0, wq[0]);
#pragma unroll
    for (int ch = 0; ch < kLeafTile / kCostLeafChunk; ++ch) {
      if (ch * kCostLeafChunk >= cnt) break;  // uniform
      if (ch + 1 < kLeafTile / kCostLeafChunk) issue(ch + 1, wq[(ch + 1) & 1]);
#pragma unroll
      for (int j = 0; j < kCostLeafChunk; ++j) {
        const int i = ch * kCostLeafChunk + j;
        if (i >= cnt) break;  // uniform
        const int32_t f = __builtin_amdgcn_readlane(flags, i);
        const int32_t w = wq[ch & 1][j] - __builtin_amdgcn_readlane(hk, i);
        const int32_t m = (f & 1) ? min(w, rprev) : w;
        int32_t r = m;
        if (f & 4) {  // pop
          r = min(m, tos);
          sp = max(sp - 1, 0);
          if (sp > 0) tos = stk[sp - 1][lane];
        }
        if ((f & 2) && sp < kCostLeafStack) {  // push
          if (sp > 0) stk[sp - 1][lane] = (int16_t)tos;
          tos = r;
          ++sp;
        }
        rprev = r;
        tile[i][lane] = (uint16_t)(__builtin_amdgcn_readlane(hsub, i) + min(dv, m));
      }
    }
    __syncthreads();
    // Row segments: line l's values at positions top - i, i in lanes.
    const bool pos_ok = lane < cnt;
    unsigned long long lm = lmask;
    while (lm) {
      const int l = __ffsll((long long)lm) - 1;
      lm &= lm - 1;
      if (pos_ok) rows[(int64_t)(v0 + l) * T.n + (top - lane)] = tile[lane][l];
    }
    // Column layout: position top - i across the tile's 64 nodes.
    const unsigned long long skip = __ballot(!pos_ok || (flags & 8));
    for (int i = 0; i < cnt; ++i) {
      if (!((skip >> i) & 1ull) && col_ok)
        cols[(int64_t)(top - i) * R.n + vl] = tile[i][lane];
    }
    __syncthreads();
  }
}

// A grid's threads (gridDim.x * blockDim.x) must stay below 2^32: launches
// over more work items than kGridThreads / threads-per-group go in pieces.
constexpr int64_t kGridThreads = int64_t{1} << 31;

void launch_leaf(const LeafArgs& a, int64_t ntiles, hipStream_t s) {
  if (a.costs) {
    constexpr int64_t kChunk = kGridThreads / 64;
    for (int64_t base = 0; base < ntiles; base += kChunk)
      hipLaunchKernelGGL(leaf_cost_kernel, dim3((unsigned)std::min(kChunk, ntiles - base)),
                         dim3(64), 0, s, a, ntiles, base);
    return;
  }
  constexpr int64_t kChunk = kGridThreads / 256;
  for (int64_t base = 0; base < ntiles; base += kChunk)
    hipLaunchKernelGGL(leaf_kernel, dim3((unsigned)std::min(kChunk, ntiles - base)), dim3(256), 0,
                       s, a, ntiles, base);
}

// ---------------------------------------------------------------------------
// Workgroup row scan: one table per workgroup of W waves, C columns per
// lane. Columns are counted from jmax: t = jmax-1-j. Wave w owns the block
// t in [w*64*C, (w+1)*64*C). Inside it, lanes own E contiguous columns per
// group and the G = C/E groups are interleaved: lane l's element (g, e) is
//   t = w*64*C + g*64*E + l*E + e,   E = min(C, 8).
// So one group of a TD row is a 64*E-int contiguous span that the wave
// reads with one coalesced E-int load per lane, and the row scan is G
// chained wave scans over groups.
// ---------------------------------------------------------------------------
// Words of a TD row group: 2, 4 or 8 entries of 16 bits (td_t), two to a
// 32-bit word, at any even byte address (the target's unaligned-access mode
// lets one dwordx2 / dwordx4 load read them).
typedef uint32_t u32x4a2 __attribute__((ext_vector_type(4), aligned(2)));
typedef uint32_t u32x2a2 __attribute__((ext_vector_type(2), aligned(2)));
typedef uint32_t u32x1a2 __attribute__((aligned(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
// Word 3 of a raw buffer resource on CDNA (32-bit data format, no swizzle).
constexpr int32_t kBufferWord3 = 0x00020000;

// ---------------------------------------------------------------------------
// Column-blocked form (kXL). A table too wide or too LDS-hungry for one
// workgroup is cut into blocks of BW = 64*W*C columns (t ranges), one
// workgroup per block, every block sweeping all rows. Per row, block b needs
// from the blocks before it
//   * the scan carry min_{t' < b*BW} (a[t'] - t') (boundary included): block
//     b-1 publishes min(its carry-in, its total); F at t = b*BW-1 is that
//     carry + b*BW-1, so the left boundary and the next row's diagonal
//     follow from it too;
//   * far values F[R][f] of saved rows R at far columns f in earlier blocks.
//     For a block those columns are the far columns of the right paths that
//     cross its right edge, one per nesting depth below the column keyroot,
//     so they live in kd_c+1 import entries of the block's slot rows. The
//     owner of f writes each such value once, at its row R, into its own
//     run of words for that row (one word per depth); every later block that
//     imports it reads it there.
// Words crossing workgroups are {epoch:32 | value:32}, written once per
// launch with device-scope relaxed atomics; a reader polls until the epoch
// matches, so no fence is needed and nothing is cleared between launches.
// Workgroups take (table, block) from a ticket counter in start order, so
// a block only ever waits on blocks that are already running. Waits are
// bounded: after kXLWaitTicks of the constant 100 MHz wall clock a
// workgroup raises the abort flag (recording what it waited for), every
// waiting workgroup then stops waiting, and the call fails.
// ---------------------------------------------------------------------------
constexpr uint64_t kXLWaitTicks = 1000000000ull;  // 10 s at 100 MHz
// A pipelined block checks its predecessor's progress every kXLEvery rows,
// one window of row metadata (64 rows).
constexpr int32_t kXLEvery = 64;
// Tables with at most xl_seq_rows rows are swept block after block by one
// workgroup (forms of at most 1024 columns): their words stay inside that
// workgroup (same CU, same L2), so they need no waits and no atomics; a
// pipeline of workgroups pays off only when rows are many.

// Writes are atomic exchanges, performed where every XCD sees the same line.
// A word is read twice at most: first with a plain load, issued rows ahead,
// which is right whenever its epoch matches (a word is written once per
// launch); a stale first read (the line can sit in the reader's XCD-local L2
// from before the write, and polls would keep hitting that copy) is polled
// again with atomic reads, which see the written value. An atomic read is a
// compare-exchange against a value no word holds (epoch 0xFFFFFFFF is never
// issued), which fails and returns the current value; an idempotent
// or-with-0 would be turned back into a plain load by the compiler.
constexpr unsigned long long kXLNever = ~0ull;
__device__ __forceinline__ unsigned long long xl_peek(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long xl_load(const unsigned long long* p) {
  unsigned long long cur = kXLNever;
  (void)__hip_atomic_compare_exchange_strong(const_cast<unsigned long long*>(p), &cur, kXLNever,
                                             __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
  return cur;
}
__device__ __forceinline__ void xl_store(unsigned long long* p, uint32_t epoch, int32_t v,
                                         bool local) {
  const unsigned long long w = ((unsigned long long)epoch << 32) | (uint32_t)v;
  if (local) __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else (void)__hip_atomic_exchange(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t xl_flag(int32_t* f) {
  int32_t cur = -1;  // the flag is 0 or 1
  (void)__hip_atomic_compare_exchange_strong(f, &cur, -1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
  return cur;
}

// Rows whose imports a column-blocked block reads together (one batch ahead).
constexpr int32_t kXLImpRows = 8;

// Lanes with `need` poll *p until it carries `epoch` (w is the first read).
// abort_flag[0] is the flag; the first workgroup to give up records
// what = (ticket, kind << 24 | row, word low, word high) in abort_flag[2..5].
__device__ __forceinline__ int32_t xl_wait(const unsigned long long* p, unsigned long long w,
                                           bool need, uint32_t epoch, int32_t* abort_flag,
                                           int32_t ticket, int32_t what) {
  uint64_t t0 = 0;
  for (int spins = 0; __ballot(need && (uint32_t)(w >> 32) != epoch); ++spins) {
    if ((spins & 63) == 0) {
      const uint64_t now = wall_clock64();
      if (spins == 0) t0 = now;
      else if (now - t0 > kXLWaitTicks) {
        const bool stale = need && (uint32_t)(w >> 32) != epoch;
        const unsigned long long first = __ballot(stale);
        if ((int)(threadIdx.x & 63) == __ffsll((long long)first) - 1 &&
            atomicCAS(abort_flag + 1, 0, 1) == 0) {
          abort_flag[2] = ticket;
          abort_flag[3] = what;
          abort_flag[4] = (int32_t)(uint32_t)w;
          abort_flag[5] = (int32_t)(w >> 32);
          abort_flag[6] = (int32_t)epoch;
        }
        (void)__hip_atomic_exchange(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (xl_flag(abort_flag)) break;
    }
    if (spins < 4) __builtin_amdgcn_s_sleep(8);
    else __builtin_amdgcn_s_sleep(32);
    if (need && (uint32_t)(w >> 32) != epoch) w = xl_load(p);
  }
  return (int32_t)(uint32_t)w;
}

// Arena words of one column-blocked task: nb-1 runs of carries (one per
// row, each run padded to whole 128-byte L2 lines), then for every saved row
// and source block a line-padded run of kds export words (by depth). A line
// thus only holds words of one producer, and no block that reads a run
// writes to it.
struct XLGeom {
  int32_t nrows, nb, nsaved, kds;
};
constexpr int32_t kXLLineWords = 16;  // 8-byte words per 128-byte L2 line
__device__ __forceinline__ int32_t xl_pad(int32_t w) {
  return (w + kXLLineWords - 1) & ~(kXLLineWords - 1);
}
__device__ __forceinline__ XLGeom xl_geom(const PoolView& P, const PairMeta& pm, uint64_t task,
                                          int bw) {
  const TreeMeta tx = P.trees[pm.tx], ty = P.trees[pm.ty];
  const int32_t i0 = (int32_t)task_i0(task), j0 = (int32_t)task_j0(task);
  const int32_t a = P.orl[tx.node_off + i0] - i0 + 1, b = P.orl[ty.node_off + j0] - j0 + 1;
  const bool rows_x = a <= b;
  const int64_t offR = rows_x ? tx.node_off : ty.node_off;
  const int64_t offC = rows_x ? ty.node_off : tx.node_off;
  const int32_t r0 = rows_x ? i0 : j0, c0 = rows_x ? j0 : i0;
  XLGeom g;
  g.nrows = rows_x ? a : b;
  g.nb = ((rows_x ? b : a) + bw - 1) / bw;
  g.nsaved = P.lrank[offR + r0 + g.nrows] - P.lrank[offR + r0];
  g.kds = P.kdn[offC + c0] + 1;
  return g;
}
__device__ __forceinline__ int64_t xl_words(const XLGeom& g) {
  return (int64_t)(g.nb - 1) * xl_pad(g.nrows) + (int64_t)g.nsaved * g.nb * xl_pad(g.kds);
}

// ---------------------------------------------------------------------------
// Workgroup row scan: one table (or, with kXL, one column block of a table)
// per workgroup of W waves, C columns per lane. Columns are counted from
// jmax: t = jmax-1-j. Wave w owns the block t in [tb0 + w*64*C,
// tb0 + (w+1)*64*C). Inside it, lanes own E contiguous columns per group and
// the G = C/E groups are interleaved: lane l's element (g, e) is
//   t = tb0 + w*64*C + g*64*E + l*E + e,   E = min(C, 8).
// So one group of a TD row is a 64*E-int contiguous span that the wave
// reads with one coalesced E-int load per lane, and the row scan is G
// chained wave scans over groups.
// ---------------------------------------------------------------------------
template <int W, int C, bool kXL, bool kCost>
__global__ __launch_bounds__(64 * W) void rowscan_kernel(DPArgs A) {
  constexpr int E = C <= 8 ? C : 8;  // contiguous columns per group
  constexpr int G = C / E;           // groups per lane
  // TD ring slots (a power of two dividing 64): up to 32 registers of ring,
  // so narrow lanes prefetch more rows ahead; 16 columns per lane take 4
  // slots as well (with 2, one row ahead, `rowscan_w1_c16` ran 2% slower at
  // 127 VGPRs against 141 with 4).
  // (The column-blocked form also rings its imports, and the cost form's
  // longer row body has to unroll S times: at most 8 slots for both.)
  constexpr int S0 = C >= 8 ? 4 : C >= 4 ? 8 : 16;
  constexpr int S = kXL && S0 > 4 ? 4 : kCost && S0 > 8 ? 8 : S0;
  constexpr int PD = S - 1;  // rows of TD prefetched
  constexpr int BW = 64 * W * C;                    // columns of one block
  constexpr int KQ = kXL ? kMaxImports / 64 : 1;    // import entries per lane
  constexpr bool kSeqForm = kXL && W * C <= 16;      // may sweep all blocks itself
  // Unit costs keep row k's values less k (F - t + T_g - k): the vertical
  // step F[r+1][c] + 1 of row k+1 is then row k's value itself, with no add
  // per column and row; the terms read from the far rows and TD take the -k
  // through their (scalar) per-column constant, and stores add k back.
  extern __shared__ uint16_t slots[];
  // Per-wave scan totals, double-buffered by row parity: a wave may write
  // row r-1's total while a slower wave still reads row r's.
  __shared__ int32_t xtot[2][W];
  __shared__ int32_t xsh[kXL ? kMaxImports + 1 : 1];  // import flags, then the ticket
  // (kXL) imports of the rows of the current batch of kXLImpRows rows, by
  // depth, and per row and quarter of depths (64 each) the lanes whose word
  // was stale. Wave w stages quarters w, w + W, ... and is their only reader.
  __shared__ uint16_t istage[kXL ? kXLImpRows : 1][kXL ? kMaxImports : 1];
  __shared__ unsigned long long istale[kXL ? kXLImpRows : 1][KQ];

  int64_t tix = blockIdx.x;
  int32_t blk = 0, ticket = 0;
  if constexpr (kXL) {
    if (threadIdx.x == 0) xsh[kMaxImports] = (int32_t)atomicAdd(A.xl_ticket, 1u);
    for (int k = threadIdx.x; k < kMaxImports; k += 64 * W) xsh[k] = 0;
    __syncthreads();
    const int32_t tk = xsh[kMaxImports];
    ticket = tk;
    tix = tk / A.xl_nbmax;
    blk = tk % A.xl_nbmax;
  }
  if (tix >= A.ntasks) return;
  const TaskDesc dsc = A.descs[tix];
  const bool rows_x = dsc.base_rx < 0;
  const int64_t offR = dsc.offR;
  const int64_t offC = dsc.offC;
  const NodeRec* __restrict__ rrecR = A.pool.rrec + offR;
  const int32_t* __restrict__ lrankR = A.pool.lrank + offR;
  const NodeRec* __restrict__ crecC = A.pool.crec + offC;
  const int32_t* __restrict__ lrankC = A.pool.lrank + offC;
  const int32_t* __restrict__ kdepC = A.pool.kdep + offC;
  const int32_t* __restrict__ rtopC = A.pool.rtop + offC;
  const int32_t r0 = dsc.r0, c0 = dsc.c0;
  td_t* TDo = A.td + dsc.tdo;  // [r][c]
  td_t* TDt = A.td + dsc.tdt;  // [c][r]
  const int64_t ldo = dsc.ldo;
  const int64_t ldt = dsc.ldt;

  const int32_t rmax = dsc.rmax, cmax = dsc.cmax;
  const int32_t ncols = cmax - c0;
  const int32_t nrows = rmax - r0;
  const int32_t base = dsc.base_rx & 0x7FFFFFFF;
  const int32_t orl_c0 = cmax - 1;
  // Non-unit costs (kCost). Hpre / Vpre: prefix sums of the horizontal
  // (column) and vertical (row) edit costs over the column and row trees:
  // insertions and deletions when the rows are x, the other way round when
  // they are y. Q(c) = Hpre[cmax-1] - Hpre[c] takes the place of t (with
  // unit costs Q(c) = t), hq = Hpre[cmax-1], hc = Hpre[cmax] (so hc - hq is
  // the cost of column cmax-1) and vr = Vpre[rmax].
  const int32_t* __restrict__ Hpre = nullptr;
  const int32_t* __restrict__ Vpre = nullptr;
  int32_t hq = 0, hc = 0, vr = 0;
  if constexpr (kCost) {
    Hpre = (rows_x ? A.pool.pins : A.pool.pdel) + offC;
    Vpre = (rows_x ? A.pool.pdel : A.pool.pins) + offR;
    hq = Hpre[cmax - 1];
    hc = Hpre[cmax];
    vr = Vpre[rmax];
  }
  int32_t nb = 1;
  bool seq = false;  // one workgroup sweeps every block of the table
  // (kXL) Wave 0's count of prefetched reads that were stale (carry or
  // import rows that had to be polled again) and of pipeline checks that had
  // to wait, added to the launch's totals at the end.
  int32_t n_stale = 0, n_wait = 0;
  if constexpr (kXL) {
    nb = (ncols + BW - 1) / BW;
    seq = kSeqForm && nrows <= A.xl_seq_rows;
    if (blk >= nb || (seq && blk > 0)) return;
  }
  do {
  if constexpr (kSeqForm) {
    if (blk > 0 && seq) {
      // The previous block's words are written; the import flags restart.
      __syncthreads();
      for (int k = threadIdx.x; k < kMaxImports; k += 64 * W) xsh[k] = 0;
      __syncthreads();
    }
  }
  const int32_t tb0 = blk * BW;  // first t of this block (0 on chip)
  // A slot row holds the compact far columns of this block (leaves but the
  // table's last, by leaf rank), at index wb F[R][cmax] = rmax - R for the
  // tree columns, and (kXL) the kds imported far values after it.
  int32_t lbase, wb, kds = 0, kdep_c0 = 0;
  if constexpr (kXL) {
    const int32_t chi = cmax - 1 - tb0;
    const int32_t clo = max(c0, cmax - tb0 - BW);
    lbase = lrankC[max(clo - 1, c0)];
    wb = max(0, lrankC[min(chi, cmax - 1)] - lbase);
    kds = A.pool.kdn[offC + c0] + 1;
    kdep_c0 = kdepC[c0];
  } else {
    lbase = dsc.lbase;
    wb = lrankC[cmax] - lbase - 1;
  }
  // The row's last entry is a sink: the saved-row stores of columns that
  // save nothing go there, so every column stores without a branch.
  const int32_t sinkix = wb + 1 + kds;
  const int32_t rowlen = sinkix + 1;

  // Uniform over the wave, so it lives in a scalar register.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int32_t tlane = tb0 + wave * 64 * C + lane * E;  // t of this lane's element (0, 0)
#define TCOL(g, e) (tlane + (g) * 64 * E + (e))

  // Per-column state, element j = g*E + e (t = TCOL(g, e), c = cmax-1-t).
  //   tcol: a tree column (orl_C(c) == orl_C(c0));
  //   lb:   on a tree column its label (tree cells only), else the far value
  //         on the boundary row F[rmax][orl_C(c)+1] = cmax - orl_C(c) - 1;
  //   foff: the column's far index in a slot row (wb for tree columns);
  //   woff: the index this column writes in a saved row (c-1 a leaf of
  //         subtree(c0)), or the sink entry;
  //   xinf: (kXL) for a written column whose right path's top keyroot lies
  //         in a later block: the path's depth index (its value is
  //         exported); else -1;
  //   Fpm:  F[r+1][c] + 1 - t + T_g (the scan works on a[t] - t);
  //   negt: (kCost) the column's -Q(c).
  // Unit costs keep the values of group g shifted by T_g = TCOL(g, 0), this
  // lane's t at the group's first element: the per-column term -t + T_g is
  // then the constant -e (an inline operand, not a register), and only the
  // wave scan of each group works unshifted. term(j) is that per-column term
  // (-Q(c) under a cost model, where nothing is shifted), gsh(g) the shift.
  // Columns past the table (t >= ncols) come after every real column in the
  // scan order, so whatever they hold never reaches a real column; only
  // their stores are masked.
  bool tcol[C];
  // tdq: the TD ring, as loaded: two 16-bit entries per word (one per word
  // when E = 1), unpacked where a row uses them (tdv).
  constexpr int NW = E == 1 ? C : C / 2;
  int32_t lb[C], foff[C], woff[C], Fpm[C];
  uint32_t tdq[S][NW];
  int32_t negt[kCost ? C : 1];
  int32_t xinf[kXL ? C : 1];
  auto term = [&](int j) -> int32_t {
    if constexpr (kCost) return negt[j];
    else return -(j % E);
  };
  auto gsh = [&](int g) -> int32_t {
    if constexpr (kCost) return 0;
    else return TCOL(g, 0);
  };
  // F[rmax][c] + (vertical cost of row rmax-1) - Q(c): 2 with unit costs.
  const int32_t fpm0 = kCost ? (hc - hq) + (vr - Vpre[rmax - 1]) : 2;
#pragma unroll
  for (int g = 0; g < G; ++g) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = g * E + e;
      const int32_t t = TCOL(g, e);
      const int32_t c = max(cmax - 1 - t, c0);
      // One record: orl(c), label, leaf rank at orl(c), leaf rank of c-1
      // (bit 31: c-1 is a leaf; only read when c > c0).
      const NodeRec cr = crecC[c];
      const int32_t oc = cr.orl;
      const int32_t lab = cr.label;
      const int32_t lro = cr.x;
      const int32_t lrm = cr.y & 0x7FFFFFFF;
      const bool prev_leaf = cr.y < 0;
      const bool tree = (oc == orl_c0);
      tcol[j] = tree;
      // (kCost) the column tree's horizontal sums at c and at orl(c) + 1
      int32_t hcol = 0, hfar = 0;
      if constexpr (kCost) {
        const CostRec cc = A.pool.ccrec[offC + c];
        hcol = rows_x ? cc.ins : cc.del;
        hfar = rows_x ? cc.ins_end : cc.del_end;
      }
      lb[j] = tree ? lab : kCost ? hc - hfar : cmax - oc - 1;
      foff[j] = tree ? wb : lro - lbase;
      const bool writes = t < ncols && c > c0 && prev_leaf;
      woff[j] = writes ? lrm - lbase : sinkix;
      // F[rmax][c] + 1 - t = cmax - c + 1 - t with unit costs, shifted
      Fpm[j] = fpm0 + gsh(g);
      if constexpr (kCost) {
        negt[j] = hcol - hq;
        // Opaque to the optimiser, so that sums with it stay three-operand
        // adds (v_add3_u32) instead of becoming an add and a subtract.
        asm("" : "+v"(negt[j]));
      }
      if constexpr (kXL) {
        // far column oc+1 in an earlier block: an import, by path depth
        const bool imp = !tree && t < ncols && cmax - 2 - oc < tb0;
        const int32_t kdc = kdepC[c] - kdep_c0;
        if (imp) {
          foff[j] = wb + 1 + kdc;
          xsh[kdc] = (cmax - 2 - oc) / BW + 1;  // the block owning the far column, plus one
        }
        const int32_t u = rtopC[max(c - 1, c0)];
        const int32_t bu = (cmax - 1 - u) / BW;
        xinf[j] = (writes && bu > blk) ? kdepC[max(c - 1, c0)] - kdep_c0 : -1;
      }
    }
  }
  // Which import entries this block uses and the block each comes from:
  // this wave's quarters of depths q = wave + W*m, lane l holding depth
  // l + 64q; hasq[m] when any lane of quarter m imports (uniform).
  constexpr int QW = (KQ + W - 1) / W;
  bool impl[QW], hasq[QW];
  int32_t isrc[QW];
  if constexpr (kXL) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < QW; ++m) {
      const int q = wave + W * m;
      const int32_t v = q < KQ ? xsh[lane + 64 * q] : 0;
      impl[m] = v != 0;
      isrc[m] = max(v - 1, 0);
      hasq[m] = __ballot(impl[m]) != 0;
    }
  }

  // TD row rr of this lane's columns into out[]: entry k of a group (column
  // off + k) is half k & 1 of word k / 2 of the group. Loads are
  // unconditional, so the compiler never waits for them inside a branch: a
  // group with no valid column reads E entries from column c0 instead, a
  // partly valid group reads up to E-1 entries before column c0 (the arena
  // keeps kTdPad entries of padding before and after every table, so both
  // stay inside the allocation).
  // On-chip forms read through a buffer resource over the table (its
  // kTdPad entries of padding included): the row offset is a scalar and each
  // lane's column offset a VGPR fixed for the whole table, so a row's loads
  // write no VGPR but their destinations (with 64-bit per-lane addresses
  // the compiler formed each row's address in the ring registers, which made
  // the loop head wait for every load in flight), and the 32-bit offsets
  // free registers (rowscan_w2_c8 80 VGPRs against 90; the 4096-pair batch
  // 99.4 against 101.3 ms). The column-blocked forms, whose rows also wait
  // on carry and import words, ran slower that way (the 20,000-node pair
  // 227 against 217 ms) and keep 64-bit addresses.
  const __amdgpu_buffer_rsrc_t tdrs =
      __builtin_amdgcn_make_buffer_rsrc((void*)(TDo - kTdPad), 0, -1, kBufferWord3);
  // Column offset of each group: in entries from the table (kXL), else in
  // bytes from the resource base.
  int32_t roff[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int32_t tb = TCOL(g, 0);
    const int32_t off = tb < ncols ? cmax - E - tb : c0;
    roff[g] = kXL ? off : (off + kTdPad) * (int32_t)sizeof(td_t);
  }
  auto load_row = [&](int32_t rr, uint32_t* out) {
    if constexpr (kXL) {
      const td_t* row = TDo + (int64_t)rr * ldo;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const td_t* p = row + roff[g];
        if constexpr (E == 8) {
          const u32x4a2 v = *reinterpret_cast<const u32x4a2*>(p);
          out[g * 4 + 0] = v.x;
          out[g * 4 + 1] = v.y;
          out[g * 4 + 2] = v.z;
          out[g * 4 + 3] = v.w;
        } else if constexpr (E == 4) {
          const u32x2a2 v = *reinterpret_cast<const u32x2a2*>(p);
          out[g * 2 + 0] = v.x;
          out[g * 2 + 1] = v.y;
        } else if constexpr (E == 2) {
          out[g] = *reinterpret_cast<const u32x1a2*>(p);
        } else {
          out[g] = *p;
        }
      }
    } else {
      // rr * ldo < 2^30 entries (trees of at most 32767 nodes): an unsigned
      // 32-bit byte offset.
      const int32_t so = (int32_t)((uint32_t)rr * (uint32_t)ldo * (uint32_t)sizeof(td_t));
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if constexpr (E == 8) {
          const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(tdrs, roff[g], so, 0);
          out[g * 4 + 0] = v.x;
          out[g * 4 + 1] = v.y;
          out[g * 4 + 2] = v.z;
          out[g * 4 + 3] = v.w;
        } else if constexpr (E == 4) {
          const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(tdrs, roff[g], so, 0);
          out[g * 2 + 0] = v.x;
          out[g * 2 + 1] = v.y;
        } else if constexpr (E == 2) {
          out[g] = __builtin_amdgcn_raw_buffer_load_b32(tdrs, roff[g], so, 0);
        } else {
          out[g] = __builtin_amdgcn_raw_buffer_load_b16(tdrs, roff[g], so, 0);
        }
      }
    }
  };
  // Element j of ring slot u as int32: column t = TCOL(j / E, j % E) is
  // entry E-1-(j % E) of its group.
  auto tdv = [&](int u, int j) -> int32_t {
    if constexpr (E == 1) {
      return (int32_t)tdq[u][j];
    } else {
      const int k = E - 1 - j % E;
      const uint32_t w = tdq[u][(j / E) * (E / 2) + k / 2];
      return (k & 1) ? (int32_t)(w >> 16) : (int32_t)(w & 0xFFFFu);
    }
  };
#pragma unroll
  for (int q = 0; q < PD; ++q) load_row(max(rmax - 1 - q, r0), tdq[q]);

  // Row metadata, 64 rows per chunk: lane i holds row rmax-1-(64q+i) as
  // (label, Rf = orl_R(r)+1, slots = fslot | wslot << 16 | wrow << 31) and,
  // for kXL, the saved row's index (leaf rank of r-1) or -1; read per row
  // with readlane; the next chunk is loaded one chunk ahead.
  auto meta_rec = [&](int q) {
    return rrecR[max(rmax - 1 - (64 * q + lane), r0)];  // rows past the table read row r0
  };
  auto load_meta = [&](int q, int32_t& mlab, int32_t& mrf, int32_t& mslot, int32_t& mrid,
                       int32_t& mvp, const NodeRec* pre = nullptr) {
    const int32_t rr = rmax - 1 - (64 * q + lane);
    const int32_t r = max(rr, r0);  // rows past the table read row r0
    const int32_t rp = max(r - 1, r0);
    const NodeRec rc = pre ? *pre : rrecR[r];  // orl(r), label, kdep(r), kdep(r-1) | r-1 a leaf
    const int32_t orr = rc.orl, lab = rc.label, kd = rc.x;
    const int32_t kdp = rc.y & 0x7FFFFFFF;
    const int32_t fs = (orr + 1 < rmax) ? kd - base - 1 : 0;
    const bool saved = r > r0 && rc.y < 0;
    const int32_t ws = saved ? (((kdp - base - 1) << 16) | (int32_t)0x80000000) : 0;
    mlab = lab;
    mrf = orr + 1;
    mslot = fs | ws;
    if constexpr (kXL) mrid = (saved && rr >= r0) ? lrankR[rp] - lrankR[r0] : -1;
    if constexpr (kCost) {
      const CostRec cr = A.pool.ccrec[offR + r];
      mvp = rows_x ? cr.del : cr.ins;  // Vpre[r]
    }
  };
  const int32_t nchunks = (nrows + 63) >> 6;
  int32_t mlab, mrf, mslot, mrid = -1, nlab = 0, nrf = 0, nslot = 0, nrid = -1;
  int32_t mvp = 0, nvp = 0;  // kCost: Vpre of the row
  {
    // The first two chunks' records are read before either is used (one
    // load latency in the set-up, not two), each as one whole record: with
    // kdep pinned the compiler does not load it in a branch of its own after
    // a wait for the rest.
    NodeRec rc0 = meta_rec(0), rc1 = meta_rec(1);
    asm volatile("" : "+v"(rc0.x), "+v"(rc1.x));
    load_meta(0, mlab, mrf, mslot, mrid, mvp, &rc0);
    if (nchunks > 1) load_meta(1, nlab, nrf, nslot, nrid, nvp, &rc1);
  }

  // Column-blocked form: the upstream carries are read a window (64 rows) at
  // a time, one window ahead, lane i holding row 64j + i's word (cwl / cwh,
  // the next window in nwl / nwh): a read that crosses XCDs takes longer
  // than a few rows, and loads return in order, so carries read a few rows
  // ahead held up every later TD load behind them. For the same reason the
  // imports of the saved rows of rows 8m .. 8m+7 are read when row 8m-8
  // starts (iw: wave w reads the quarters of depths q = w, w + W, ..., lane
  // l depth l + 64q) and moved into the LDS stage when row 8m starts; a
  // word that is stale is polled again when its row uses it.
  unsigned long long* xw = A.xl_words;
  int64_t cin_off = 0, cout_off = 0, exp_off = 0;
  const int32_t kdsp = xl_pad(kds);
  if constexpr (kXL) {
    const int64_t cstride = xl_pad(nrows);
    xw += A.xl_off[tix];
    cin_off = (int64_t)(blk - 1) * cstride;
    cout_off = (int64_t)blk * cstride;
    exp_off = (int64_t)(nb - 1) * cstride;
  }
  uint32_t cwl = 0, cwh = 0, nwl = 0, nwh = 0;
  unsigned long long iw[QW][kXL ? kXLImpRows : 1];
  auto load_cwin = [&](int32_t w0, uint32_t& lo, uint32_t& hi) {
    if constexpr (kXL) {
      const int32_t row = w0 + lane;
      if (blk > 0 && row < nrows) {
        const unsigned long long w = xl_peek(xw + cin_off + row);
        lo = (uint32_t)w;
        hi = (uint32_t)(w >> 32);
      }
    }
  };
  // Export word of saved row rid written by block bb at depth index kx.
  auto xword = [&](int32_t rid, int32_t bb, int32_t kx) {
    return xw + exp_off + ((int64_t)rid * nb + bb) * kdsp + kx;
  };
  // Rows q0 .. q0+7 (rows k0m .. k0m+63 have their metadata in mrid, the
  // next 64 in nrid): each wave reads its quarters' imports of their saved
  // rows into iw.
  auto iq_issue = [&](int32_t q0, int32_t k0m) {
    if constexpr (kXL) {
#pragma unroll
      for (int m = 0; m < QW; ++m) {
        if (!hasq[m]) continue;  // uniform
#pragma unroll
        for (int r = 0; r < kXLImpRows; ++r) {
          const int32_t k = q0 + r;
          const int32_t ix = k - k0m;  // uniform, in [0, 128)
          const int32_t rid = k >= nrows ? -1
                              : ix < 64 ? __builtin_amdgcn_readlane(mrid, ix)
                                        : __builtin_amdgcn_readlane(nrid, ix - 64);
          iw[m][r] = (unsigned long long)A.xl_epoch << 32;  // nothing to read: valid
          if (rid >= 0 && impl[m])
            iw[m][r] = xl_peek(xword(rid, isrc[m], lane + 64 * (wave + W * m)));
        }
      }
    }
  };
  // The batch read last time becomes the stage of the rows starting now.
  auto iq_land = [&]() {
    if constexpr (kXL) {
#pragma unroll
      for (int m = 0; m < QW; ++m) {
        if (!hasq[m]) continue;  // uniform
        const int q = wave + W * m;
#pragma unroll
        for (int r = 0; r < kXLImpRows; ++r) {
          const bool stale = (uint32_t)(iw[m][r] >> 32) != A.xl_epoch;
          istage[r][lane + 64 * q] = (uint16_t)(uint32_t)iw[m][r];
          const unsigned long long sm = __ballot(stale);
          if (lane == 0) istale[r][q] = sm;
        }
      }
    }
  };
  // At row k = 64j, before the carries of window j+1 (rows up to 64j+127)
  // and the imports of rows up to k+PD+63 are read, the block before this
  // one has finished row 64j+127 (the end of a line), and so have the
  // exporters, which run ahead of it: those reads find their lines complete.
  // One atomic poll per kXLEvery rows, read half a window ahead (tw) so its
  // round trip overlaps those rows.
  unsigned long long tw = 0;
  auto xl_throttle_row = [&](int32_t k) { return min(k + 2 * kXLEvery - 1, nrows - 1); };
  auto xl_throttle_word = [&](int32_t k) { return xw + cin_off + xl_throttle_row(k); };
  auto xl_throttle = [&](int32_t k, bool early) {
    if constexpr (kXL) {
      if (blk > 0 && !seq) {
        const int32_t t = xl_throttle_row(k);
        const unsigned long long w0 = early ? tw : xl_load(xl_throttle_word(k));
        n_wait += (uint32_t)(w0 >> 32) != A.xl_epoch;
        (void)xl_wait(xl_throttle_word(k), w0, true, A.xl_epoch, A.xl_abort, ticket, t);
      }
    }
  };
  if constexpr (kXL) {
    xl_throttle(0, false);
    load_cwin(0, cwl, cwh);
    load_cwin(kXLEvery, nwl, nwh);
    iq_issue(0, 0);
    iq_land();
    iq_issue(kXLImpRows, 0);
  }

  // F[r+1] at the column just before this wave's block (t = tb0 + w*64*C - 1)
  // for the row about to start: the boundary row's t + 1 first, then each
  // row's carry-in plus that t.
  // (With costs: Q at that column, qleft, and F[rmax] there.)
  const int32_t cleft = max(cmax - tb0 - wave * 64 * C, c0);
  const int32_t qleft = kCost ? hq - Hpre[cleft] : tb0 + wave * 64 * C - 1;
  int32_t left_prev = kCost ? hc - Hpre[cleft] : tb0 + wave * 64 * C;
  int32_t vp_prev = vr;  // kCost: Vpre of the previous row

  // Rows go S at a time in one flat loop with the inner loop fully unrolled,
  // so row k of the table always uses ring slot k % S and every slot index is
  // static: row k reads slot u and prefetches row k+PD into slot (u+PD) % S,
  // the slot the previous row has just finished with. Only the last
  // iteration has rows past the table's end: rings of at most 4 slots still
  // issue those rows' prefetches and skip their bodies, longer rings leave the
  // unrolled rows early (see the loop body).
  for (int32_t k0 = 0; k0 < nrows; k0 += S) {
    if ((k0 & 63) == 0 && k0 > 0) {  // next 64-row chunk of row metadata
      mlab = nlab;
      mrf = nrf;
      mslot = nslot;
      mrid = nrid;
      mvp = nvp;
      if ((k0 >> 6) + 1 < nchunks) load_meta((k0 >> 6) + 1, nlab, nrf, nslot, nrid, nvp);
    }
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const int32_t k = k0 + u;  // rows done so far
      // With a ring of at most 4 slots every slot issues its prefetch, also
      // past the last row, so that every path into the loop head has issued
      // the same loads and no wait there drains the ring (longer rings keep
      // the early exit: short tables would step through up to 15 idle slots).
      if constexpr (S <= 4) {
        load_row(max(rmax - 1 - k - PD, r0), tdq[(u + PD) % S]);
        if (k >= nrows) continue;
      } else {
        if (k >= nrows) break;
        load_row(max(rmax - 1 - k - PD, r0), tdq[(u + PD) % S]);
      }
      const int32_t i = k & 63;  // index in the metadata chunk
      const int32_t r = rmax - 1 - k;
      const int32_t ku = kCost ? 0 : k;  // the unit form's row shift
      if constexpr (kXL) {
        constexpr int32_t kM = kXLEvery - 1, kHalf = kXLEvery / 2;
        if ((k & kM) == 0 && k > 0) {
          cwl = nwl;
          cwh = nwh;
          xl_throttle(k, true);
          load_cwin(k + kXLEvery, nwl, nwh);
        }
        if ((k & kM) == kHalf && blk > 0 && !seq) tw = xl_load(xl_throttle_word(k + kHalf));
        if ((k & (kXLImpRows - 1)) == 0 && k > 0) {
          iq_land();
          iq_issue(k + kXLImpRows, k & ~63);
        }
      }

      const int32_t labr = __builtin_amdgcn_readlane(mlab, i);
      const int32_t Rf = __builtin_amdgcn_readlane(mrf, i);
      const int32_t sl = __builtin_amdgcn_readlane(mslot, i);
      const bool tree_row = (Rf == rmax);
      // Vertical costs of this row and of the next, and F[r][cmax] (unit
      // costs: 1, 1 and rmax - r).
      int32_t vcur = 1, vnext = 1, frc = rmax - r;
      if constexpr (kCost) {
        const int32_t vpr = __builtin_amdgcn_readlane(mvp, i);
        const int32_t vpn = i < 63 ? __builtin_amdgcn_readlane(mvp, i + 1)
                                   : __builtin_amdgcn_readlane(nvp, 0);
        vcur = vp_prev - vpr;
        vnext = vpr - vpn;
        frc = vr - vpr;
        vp_prev = vpr;
      }
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
        // Tree rows: F[r+1][c+1] of each group's element e = 0, the previous
        // row's value one column to the left in t order (F = Fpm - 1 + t).
        int32_t fpr0[G];
        if constexpr (kTree) {
#pragma unroll
          for (int g = 0; g < G; ++g) {
            const int jl = g * E + E - 1;  // the group's last column
            const int32_t fill =
                g ? __builtin_amdgcn_readlane(Fpm[jl - E] - vcur - term(jl - E), 63) : left_prev;
            fpr0[g] = wave_shr1(Fpm[jl] - vcur - term(jl), fill);
          }
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
              // Blended with a mask rather than a branch per element.
              int32_t fprm, rename;
              if constexpr (kCost) {
                fprm = e ? Fpm[j - 1] - vcur - term(j - 1) + term(j) : fpr0[g] + term(j);
                // Columns off the tree path read entry 0 (one line for the
                // wave) instead of a row of the matrix.
                rename = A.ren[tcol[j] ? (rows_x ? labr * A.alphabet + lb[j]
                                                 : lb[j] * A.alphabet + labr)
                                       : 0];
              } else {
                fprm = e ? Fpm[j - 1] - 2 : fpr0[g] + term(j);
                rename = (int32_t)(labr != lb[j]);
              }
              const int32_t diag = fprm + rename;
              const int32_t far = lb[j] + tdv(u, j) + (term(j) - ku);
              // Blended rather than branched per element.
              dm = tcol[j] ? diag : far;
            } else {
              dm = sv[j] + tdv(u, j) + (term(j) - ku);
            }
            h[j] = min(dm, Fpm[j]);
            if (e) h[j] = min(h[j], h[j - 1]);
          }
          const int32_t incl = wave_incl_min(h[g * E + E - 1] - gsh(g));
          // (values stay far below kBig, so group 0 needs no min with it)
          pre[g] = g ? min(carry_w, wave_shr1(incl, kBig)) : wave_shr1(incl, kBig);
          carry_w = g ? min(carry_w, __builtin_amdgcn_readlane(incl, 63))
                      : __builtin_amdgcn_readlane(incl, 63);
        }
      };
      if (tree_row) row_terms(std::true_type{});
      else row_terms(std::false_type{});

      // Carry into this wave: the boundary F[r][cmax] - (-1), or for a
      // later column block the carry of the blocks before it, then the
      // totals of this workgroup's earlier waves.
      int32_t carry = kCost ? frc + (hc - hq) : 2;  // unit: (rmax - r) + 1 - k
      if constexpr (kXL) {
        if (blk > 0) {
          const uint32_t clo = __builtin_amdgcn_readlane(cwl, i);
          if (__builtin_amdgcn_readlane(cwh, i) == A.xl_epoch) {
            carry = (int32_t)clo;
          } else {
            n_stale += wave == 0;
            carry = xl_wait(xw + cin_off + k, 0ull, true, A.xl_epoch, A.xl_abort, ticket, k);
          }
        }
      }
      if constexpr (W > 1) {
        int32_t* xt = xtot[r & 1];
        if (lane == 0) xt[wave] = carry_w;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < W - 1; ++w) carry = min(carry, w < wave ? xt[w] : kBig);
      }
      if constexpr (kXL) {
        // The next block's carry: this block's last wave, carry-in
        // combined with its own total.
        if (blk + 1 < nb && wave == W - 1 && lane == 0)
          xl_store(xw + cout_off + k, A.xl_epoch, min(carry, carry_w), seq);
      }
      // F[r][c] at the column just before this wave's block: the previous
      // wave's last element ends with h = min(its carry-in, its own total) =
      // this wave's carry-in, and F = h + t; for wave 0 of block 0 this gives
      // F[r][cmax] = rmax - r.
      left_prev = carry + qleft - (kCost ? 0 : 1);  // unit: in row k+1's shift
      // h becomes G - t; Fpm = G + 1 - t = h + 1 (unit costs: row k+1 is shifted
      // by one more, so Fpm = h).
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int32_t cg = min(carry, pre[g]) + gsh(g);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int j = g * E + e;
          h[j] = min(cg, h[j]);
          Fpm[j] = kCost ? h[j] + vnext : h[j];
        }
      }

      // Tree cells -> TD / TD^T (tree rows only); saved rows -> LDS slots.
      if (tree_row) {
        // Column of element (0, 0), made opaque here so that the compiler
        // forms these rare stores' addresses when they happen instead of
        // keeping two 64-bit addresses per element live across the loop.
        int32_t cb = cmax - 1 - tlane;
        asm volatile("" : "+v"(cb));
        td_t* const rowp = TDo + (int64_t)r * ldo;
        td_t* const colp = TDt + r;
        const uint32_t ldt32 = (uint32_t)ldt;
#pragma unroll
        for (int j = 0; j < C; ++j) {
          const int32_t c = cb - ((j / E) * 64 * E + j % E);
          if (c >= c0 && tcol[j]) {
            // c * ldt < 2^30: trees have at most 32767 nodes
            rowp[c] = (td_t)(h[j] - term(j) + ku);
            colp[(uint32_t)c * ldt32] = (td_t)(h[j] - term(j) + ku);
          }
        }
      }
      if (wrow) {
        // Columns this lane does not save write the row's sink entry.
#pragma unroll
        for (int j = 0; j < C; ++j) wrow_ptr[woff[j]] = (uint16_t)(h[j] - term(j) + ku);
        if (threadIdx.x == 0) wrow_ptr[wb] = (uint16_t)frc;
        if constexpr (kXL) {
          // Imported far values of this saved row; far values later
          // blocks import from this one.
          const int32_t rid = __builtin_amdgcn_readlane(mrid, i);
          const int ri = k & (kXLImpRows - 1);
#pragma unroll
          for (int m = 0; m < QW; ++m) {
            if (!hasq[m]) continue;  // uniform
            const int q = wave + W * m;
            const unsigned long long sm = istale[ri][q];  // uniform
            int32_t val = istage[ri][lane + 64 * q];
            if (sm) {
              ++n_stale;
              const bool stale = (sm >> lane) & 1ull;
              const int32_t v = xl_wait(xword(rid, isrc[m], lane + 64 * q), 0ull, stale,
                                        A.xl_epoch, A.xl_abort, ticket, (1 << 24) | k);
              if (stale) val = v;
            }
            if (impl[m]) wrow_ptr[wb + 1 + lane + 64 * q] = (uint16_t)val;
          }
#pragma unroll
          for (int j = 0; j < C; ++j) {
            if (xinf[j] >= 0)
              xl_store(xword(rid, blk, xinf[j]), A.xl_epoch, h[j] - term(j) + ku, seq);
          }
        }
        // The next row may read this slot row at columns of other waves.
        // (Every other LDS reuse is ordered by the barrier above: a slot
        // row is only rewritten after the rows that read it.)
        if constexpr (W > 1) __syncthreads();
      }
    }
  }
  } while (kSeqForm && seq && ++blk < nb);  // blocks of a table swept by one workgroup
  if constexpr (kXL) {
    // Stale reads: carries counted by wave 0, imports by the wave that
    // stages them; pipeline waits by wave 0 (every wave waits alike).
    if ((threadIdx.x & 63) == 0 && n_stale) atomicAdd(A.xl_abort + 8, n_stale);
    if (threadIdx.x == 0 && n_wait) atomicAdd(A.xl_abort + 9, n_wait);
  }
#undef TCOL
}

// ---------------------------------------------------------------------------
// Tiny tables (at most kSeg = 16, 32 or 64 columns, so at most kSeg rows):
// 64 / kSeg per wave, one per kSeg-lane segment, one column per lane
// (t = lane % kSeg). A segment is one or more whole 16-lane DPP rows, so the
// row's scan and shift stay in registers (row_shr, row_bcast, wave_shr) and a
// table needs no barrier and no LDS exchange. Row metadata sits in lanes:
// lane t of a table holds that of rows t % 16 + 16 s in set s (each 16-lane
// row of a segment holds every set), and row k takes set k / 16 with
// row_newbcast:(k % 16). The TD values of all rows are read up front. Tables
// of one wave run to the longest one's row count, each lane storing only
// while its table has rows left.
// ---------------------------------------------------------------------------
template <class F, int... K>
__device__ __forceinline__ void static_for(F& f, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>{}), ...);
}

#define DPP_ROW_NEWBCAST(n) (0x150 + (n))

// Inclusive prefix minimum over each kSeg-lane segment (lane order).
template <int kSeg>
__device__ __forceinline__ int seg_incl_min(int v) {
  if constexpr (kSeg == 64) {
    return wave_incl_min(v);
  } else {
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_SHR(1), 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_SHR(2), 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_SHR(4), 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_SHR(8), 0xf, 0xf, false));
    if constexpr (kSeg == 32)
      v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, DPP_ROW_BCAST15, 0xa, 0xf, false));
    return v;
  }
}

// Value of the previous lane of the segment (its lane t = 0 receives fill).
template <int kSeg>
__device__ __forceinline__ int seg_shr1(int v, int fill, int t) {
  if constexpr (kSeg == 16) {
    return __builtin_amdgcn_update_dpp(fill, v, DPP_ROW_SHR(1), 0xf, 0xf, false);
  } else {
    const int u = wave_shr1(v, fill);
    return kSeg == 64 || t != 0 ? u : fill;
  }
}

// Value of set k / 16 in lane k % 16 of this lane's 16-lane row.
template <int k, int kSets>
__device__ __forceinline__ int32_t row_meta(const int32_t (&m)[kSets]) {
  return __builtin_amdgcn_update_dpp(0, m[k / 16], DPP_ROW_NEWBCAST(k % 16), 0xf, 0xf, false);
}

// kCost: the cost form, with the substitutions of the row scan (Q(c) from
// the column tree's prefix sums in place of t, the row node's vertical cost
// in place of 1, a rename gather on tree cells).
template <int kSeg, bool kCost>
__global__ __launch_bounds__(256) void tiny_kernel(DPArgs A, int32_t slot_ints) {
  constexpr int kPerGroup = 256 / kSeg;  // tables per workgroup of 4 waves
  constexpr int kSets = kSeg / 16;       // metadata registers per lane
  extern __shared__ uint16_t tslots[];
  __shared__ uint16_t sink[256];
  const int t = threadIdx.x & (kSeg - 1);  // this lane's column, t = cmax-1-c
  const int seg = threadIdx.x / kSeg;      // table of this lane in the workgroup
  const int64_t tix = (int64_t)blockIdx.x * kPerGroup + seg;
  const bool live = tix < A.ntasks;
  const TaskDesc dsc = A.descs[live ? tix : 0];  // idle segments redo task 0, storing nothing
  const int64_t offR = dsc.offR;
  const int64_t offC = dsc.offC;
  const NodeRec* __restrict__ rrecR = A.pool.rrec + offR;
  const NodeRec* __restrict__ crecC = A.pool.crec + offC;
  const int32_t* __restrict__ lrankC = A.pool.lrank + offC;
  const int32_t r0 = dsc.r0, c0 = dsc.c0;
  td_t* TDo = A.td + dsc.tdo;  // [r][c]
  td_t* TDt = A.td + dsc.tdt;  // [c][r]
  const int64_t ldo = dsc.ldo;
  const int64_t ldt = dsc.ldt;
  const int32_t rmax = dsc.rmax, cmax = dsc.cmax;
  const int32_t ncols = cmax - c0, nrows = rmax - r0;
  const int32_t base = dsc.base_rx & 0x7FFFFFFF;
  const int32_t lbase = dsc.lbase;
  const int32_t wb = lrankC[cmax] - lbase - 1;
  const bool rows_x = dsc.base_rx < 0;
  // Cost prefix sums of the column (Hpre) and row (Vpre) trees; hq, hc, vr
  // as in the row scan.
  const int32_t* __restrict__ Hpre = nullptr;
  const int32_t* __restrict__ Vpre = nullptr;
  int32_t hq = 0, hc = 0, vr = 0;
  if constexpr (kCost) {
    Hpre = (rows_x ? A.pool.pins : A.pool.pdel) + offC;
    Vpre = (rows_x ? A.pool.pdel : A.pool.pins) + offR;
    hq = Hpre[cmax - 1];
    hc = Hpre[cmax];
    vr = Vpre[rmax];
  }
  const int32_t rowlen = wb + 1;
  uint16_t* slots = tslots + seg * (2 * slot_ints);

  // This lane's column (as in the row scan, one element).
  const int32_t c = max(cmax - 1 - t, c0);
  const NodeRec cr = crecC[c];  // orl(c), label, leaf rank at orl(c), leaf rank of c-1 | c-1 a leaf
  const int32_t oc = cr.orl;
  const int32_t labc = cr.label;
  const bool tree_col = oc == cmax - 1;
  int32_t hcol = 0, hfar = 0;  // (kCost) Hpre[c], Hpre[oc + 1]
  if constexpr (kCost) {
    const CostRec cc = A.pool.ccrec[offC + c];
    hcol = rows_x ? cc.ins : cc.del;
    hfar = rows_x ? cc.ins_end : cc.del_end;
  }
  const int32_t bfar = tree_col ? -1 : kCost ? hc - hfar : cmax - oc - 1;
  const int32_t foff = tree_col ? wb : cr.x - lbase;
  const bool col_ok = live && t < ncols;
  const bool writes = col_ok && c > c0 && cr.y < 0;
  uint16_t* const wdst_rel = writes ? slots + ((cr.y & 0x7FFFFFFF) - lbase) : sink + threadIdx.x;
  int32_t negt = kCost ? hcol - hq : -t;
  asm("" : "+v"(negt));

  // Metadata of rows t % 16 + 16 s (rows past the table read row r0).
  int32_t mlab[kSets], mrf[kSets], mslot[kSets], mvp[kSets];
#pragma unroll
  for (int s = 0; s < kSets; ++s) {
    const int32_t r = max(rmax - 1 - ((t & 15) + 16 * s), r0);
    const NodeRec rc = rrecR[r];  // orl(r), label, kdep(r), kdep(r-1) | r-1 a leaf
    const int32_t orr = rc.orl, kd = rc.x;
    const int32_t kdp = rc.y & 0x7FFFFFFF;
    const int32_t fs = (orr + 1 < rmax) ? kd - base - 1 : 0;
    const bool saved = r > r0 && rc.y < 0;
    mlab[s] = rc.label;
    mrf[s] = orr + 1;
    mslot[s] = fs | (saved ? (((kdp - base - 1) << 16) | (int32_t)0x80000000) : 0);
    mvp[s] = 0;
    if constexpr (kCost) {
      const CostRec vc = A.pool.ccrec[offR + r];
      mvp[s] = rows_x ? vc.del : vc.ins;  // Vpre[r]
    }
  }
  // TD of this lane's column for every row (row k at r = rmax-1-k).
  int32_t tdv[kSeg];
#pragma unroll
  for (int k = 0; k < kSeg; ++k)
    tdv[k] = k < nrows ? TDo[(int64_t)(rmax - 1 - k) * ldo + c] : 0;

  // Rows of the longest table of this wave.
  int32_t nmax = __builtin_amdgcn_readlane(nrows, 0);
#pragma unroll
  for (int s = 1; s < 64 / kSeg; ++s) nmax = max(nmax, __builtin_amdgcn_readlane(nrows, s * kSeg));

  // F[r+1][c] + (vertical cost of row r) - Q(c): 2 with unit costs.
  int32_t Fpm = kCost ? (hc - hq) + (vr - Vpre[rmax - 1]) : 2;
  int32_t left_prev = 0;    // F[r+1][cmax]
  int32_t vp_prev = vr;     // kCost: Vpre of the previous row
  // Row k (a compile-time index: it selects the DPP broadcast lane).
  auto row = [&](auto kc) {
    constexpr int k = decltype(kc)::value;
    if (k >= nmax) return;  // uniform
    const bool act = k < nrows;
    const int32_t r = rmax - 1 - k;
    const int32_t labr = row_meta<k>(mlab);
    const int32_t Rf = row_meta<k>(mrf);
    const int32_t sl = row_meta<k>(mslot);
    const bool tree_row = Rf == rmax;
    // Vertical costs of this row and of the next, and F[r][cmax] (unit
    // costs: 1, 1 and rmax - r).
    int32_t vcur = 1, vnext = 1, frc = rmax - r;
    if constexpr (kCost) {
      const int32_t vpr = row_meta<k>(mvp);
      const int32_t vpn = row_meta<k + 1 < kSeg ? k + 1 : k>(mvp);
      vcur = vp_prev - vpr;
      vnext = vpr - vpn;
      frc = vr - vpr;
      vp_prev = vpr;
    }
    // F[r+1][c+1] - Q(c+1): the left neighbour's previous F (lane t-1), or
    // the boundary F[r+1][cmax] for t = 0.
    const int32_t fpr = seg_shr1<kSeg>(Fpm - vcur - negt, left_prev, t);
    int32_t rename;
    if constexpr (kCost)
      rename = A.ren[(rows_x ? labr * A.alphabet + labc : labc * A.alphabet + labr) &
                     (tree_col ? -1 : 0)];
    else
      rename = (int32_t)(labr != labc);
    const int32_t diag = fpr + negt + rename;
    const int32_t sv = slots[(sl & 0xFFFF) * rowlen + foff];
    const int32_t far = (tree_row ? bfar : sv) + tdv[k] + negt;
    const int32_t dm = tree_row && tree_col ? diag : far;
    const int32_t carry = kCost ? frc + (hc - hq) : rmax - r + 1;
    const int32_t h = min(carry, seg_incl_min<kSeg>(min(dm, Fpm)));  // F - Q(c)
    if (act) {
      Fpm = h + vnext;
      left_prev = frc;
      if (tree_row && tree_col && col_ok) {
        TDo[(int64_t)r * ldo + c] = (td_t)(h - negt);
        TDt[(int64_t)c * ldt + r] = (td_t)(h - negt);
      }
      if (sl < 0) {
        uint16_t* wrow = slots + ((sl >> 16) & 0x7FFF) * rowlen;
        *(writes ? wdst_rel + (wrow - slots) : wdst_rel) = (uint16_t)(h - negt);
        if (t == 0) wrow[wb] = (uint16_t)frc;
      }
    }
  };
  static_for(row, std::make_integer_sequence<int, kSeg>{});
}

template <int kSeg>
void launch_tiny_seg(const DPArgs& a, int slot_bytes, hipStream_t s) {
  constexpr int kPerGroup = 256 / kSeg;
  const int slot_ints = (slot_bytes + 3) / 4;  // per table, in 4-byte units
  constexpr int64_t kChunk = kGridThreads / 256 * kPerGroup;
  for (int64_t off = 0; off < a.ntasks; off += kChunk) {
    DPArgs c = a;
    c.tasks = a.tasks + off;
    c.descs = a.descs + off;
    c.ntasks = std::min(kChunk, a.ntasks - off);
    const dim3 grid((unsigned)((c.ntasks + kPerGroup - 1) / kPerGroup));
    const size_t lds = (size_t)kPerGroup * slot_ints * 4;
    if (a.costs)
      hipLaunchKernelGGL((tiny_kernel<kSeg, true>), grid, dim3(256), lds, s, c, slot_ints);
    else
      hipLaunchKernelGGL((tiny_kernel<kSeg, false>), grid, dim3(256), lds, s, c, slot_ints);
  }
}

void launch_tiny(int form, const DPArgs& a, int slot_bytes, hipStream_t s) {
  switch (form) {
    case 0: launch_tiny_seg<16>(a, slot_bytes, s); break;
    case 1: launch_tiny_seg<32>(a, slot_bytes, s); break;
    default: launch_tiny_seg<64>(a, slot_bytes, s); break;
  }
}

// Arena offsets of every column-blocked task of a batch, in bucket order
// (level, then form), and their total (one workgroup).
__global__ __launch_bounds__(1024) void xl_layout_kernel(XLLayoutArgs A) {
  __shared__ unsigned long long part[1024];
  __shared__ unsigned long long run;
  const int tid = threadIdx.x;
  if (tid == 0) run = 0;
  __syncthreads();
  int64_t ord = 0;
  for (int L = 0; L < A.nlevels; ++L) {
    for (int k = 0; k < kNumXLForms; ++k) {
      const int bkt = bucket_of(L, kFirstXLClass + k, 0);
      const int64_t cnt = (int64_t)A.counts[bkt];
      const int64_t start = (int64_t)A.cursor[bkt] - cnt;  // cursors end past the bucket
      for (int64_t c0 = 0; c0 < cnt; c0 += 1024) {
        const int64_t idx = c0 + tid;
        unsigned long long w = 0;
        if (idx < cnt) {
          const uint64_t task = A.tasks[start + idx];
          w = (unsigned long long)xl_words(xl_geom(A.pool, A.pairs[task_pair(task)], task,
                                                   xl_bw(k)));
        }
        part[tid] = w;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
          const unsigned long long v = tid >= off ? part[tid - off] : 0;
          __syncthreads();
          part[tid] += v;
          __syncthreads();
        }
        if (idx < cnt) A.off[ord + idx] = (int64_t)(run + part[tid] - w);
        __syncthreads();
        if (tid == 1023) run += part[1023];
        __syncthreads();
      }
      ord += cnt;
    }
  }
  if (tid == 0) *A.total = run;
}

void launch_xl_layout(const XLLayoutArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(xl_layout_kernel, dim3(1), dim3(1024), 0, s, a);
}

// Zero fill for a newly grown column-blocked arena (epoch 0 is never used).
__global__ void xl_zero_kernel(unsigned long long* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
void launch_xl_zero(unsigned long long* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(xl_zero_kernel, dim3(1024), dim3(256), 0, s, p, n);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int W, int C>
static void launch_row(const DPArgs& a, int lds_bytes, hipStream_t s) {
  // One table per workgroup: a long task list goes in pieces.
  constexpr int64_t kChunk = kGridThreads / (64 * W);
  for (int64_t off = 0; off < a.ntasks; off += kChunk) {
    DPArgs c = a;
    c.tasks = a.tasks + off;
    c.descs = a.descs + off;
    c.ntasks = std::min(kChunk, a.ntasks - off);
    if (a.costs)
      hipLaunchKernelGGL((rowscan_kernel<W, C, false, true>), dim3((unsigned)c.ntasks),
                         dim3(64 * W), (size_t)lds_bytes, s, c);
    else
      hipLaunchKernelGGL((rowscan_kernel<W, C, false, false>), dim3((unsigned)c.ntasks),
                         dim3(64 * W), (size_t)lds_bytes, s, c);
  }
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

template <int W, int C>
static void launch_xl(const DPArgs& a, int lds_bytes, hipStream_t s) {
  // Workgroups take (task, block) tickets from one counter that the pieces
  // of a long launch share: a later piece continues where the earlier one
  // stopped, and every block it waits on holds a lower ticket, taken by a
  // workgroup of an earlier (finished) piece or a running one of its own.
  constexpr int64_t kChunk = kGridThreads / (64 * W);
  const int64_t total = a.ntasks * a.xl_nbmax;
  for (int64_t off = 0; off < total; off += kChunk) {
    const dim3 grid((unsigned)std::min(kChunk, total - off));
    if (a.costs)
      hipLaunchKernelGGL((rowscan_kernel<W, C, true, true>), grid, dim3(64 * W),
                         (size_t)lds_bytes, s, a);
    else
      hipLaunchKernelGGL((rowscan_kernel<W, C, true, false>), grid, dim3(64 * W),
                         (size_t)lds_bytes, s, a);
  }
}

// Grants every row-scan instantiation the LDS it may ask for (> 64 KiB needs
// an explicit attribute).
template <int W, int C, bool kXL = false>
static hipError_t allow_lds(int limit) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&rowscan_kernel<W, C, kXL, false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, limit);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&rowscan_kernel<W, C, kXL, true>),
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
  if ((e = allow_lds_c<16>(limit)) != hipSuccess) return e;
  // column-blocked forms (xl_waves / xl_cpl)
  if ((e = allow_lds<8, 8, true>(limit)) != hipSuccess) return e;
  if ((e = allow_lds<8, 4, true>(limit)) != hipSuccess) return e;
  if ((e = allow_lds<4, 4, true>(limit)) != hipSuccess) return e;
  if ((e = allow_lds<2, 4, true>(limit)) != hipSuccess) return e;
  if ((e = allow_lds<2, 2, true>(limit)) != hipSuccess) return e;
  return allow_lds<2, 1, true>(limit);
}

// Launches one (class, task list). Returns false for a class without a kernel.
bool launch_forest(int cls, const DPArgs& a, int lds_bytes, hipStream_t s) {
  if (a.ntasks <= 0) return true;
  if (is_tiny(cls)) {
    launch_tiny(cls - kClassTiny, a, lds_bytes, s);
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
  switch (cls - kFirstXLClass) {  // must match xl_waves / xl_cpl
    case 0: launch_xl<8, 8>(a, lds_bytes, s); return true;
    case 1: launch_xl<8, 4>(a, lds_bytes, s); return true;
    case 2: launch_xl<4, 4>(a, lds_bytes, s); return true;
    case 3: launch_xl<2, 4>(a, lds_bytes, s); return true;
    case 4: launch_xl<2, 2>(a, lds_bytes, s); return true;
    case 5: launch_xl<2, 1>(a, lds_bytes, s); return true;
    default: return false;
  }
}

}  // namespace xted
