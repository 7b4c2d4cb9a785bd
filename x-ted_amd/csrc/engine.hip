// Host engine: tree pool management, batch planning and the level loop.
//
// Call sequence of run_batch (the reference's equivalent is the per-level
// host loop of X-TED_GPU/parallel_X_TED.cu:169-399):
//   1. pair table + plan-block offsets to the device (one copy each)
//   2. plan kernels (count / scan / scatter) build every (level, class)
//      bucket for every pair of the batch
//   3. one read-back of the bucket counts
//   4. level L = 0..Lmax: one launch per non-empty class bucket; kernel
//      boundaries order the levels, so no host sync between levels
//   5. distances gathered on the device and copied back (optionally TDs)
#include <algorithm>
#include <chrono>
#include <cstring>

#include "engine.hpp"
#include "kernels.hpp"

namespace xted {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));         \
      return ERR_DEVICE;                                                           \
    }                                                                              \
  } while (0)

#define ST_TRY(expr)            \
  do {                          \
    int st_ = (expr);           \
    if (st_ != OK) return st_;  \
  } while (0)

// Trees are limited so that forest values fit the uint16 LDS slots
// (F <= |subtree x| + |subtree y|) and node ids fit the task encoding.
constexpr int32_t kMaxTreeNodes = 32767;

int DevBuf::grow(size_t bytes, bool keep, hipStream_t s) {
  if (bytes <= cap) return OK;
  size_t ncap = std::max(bytes, cap + cap / 2);
  ncap = (ncap + 4095) & ~(size_t)4095;
  void* np = nullptr;
  hipError_t e = hipMalloc(&np, ncap);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("device allocation of " + std::to_string(ncap) + " bytes failed: " +
              hipGetErrorString(e));
    return ERR_OUT_OF_MEMORY;
  }
  if (p) {
    if (keep && cap) HIP_TRY(hipMemcpyAsync(np, p, cap, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipFree(p));
  }
  p = np;
  cap = ncap;
  return OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

Context::Context(int device) : device_(device) {}

Context::~Context() {
  if (stream_) {
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(stream_);
  }
  for (DevBuf* b : {&parent_, &label_, &orl_, &kdep_, &kdn_, &lrank_, &kr_, &krkd_, &krsize_,
                    &krleaf_, &trees_, &prep_scratch_, &prep_off_, &pairs_, &blk_, &counts_,
                    &cells_, &cursor_, &maxlds_, &tasks_, &td_, &dist_, &skr_, &skd_, &ssize_,
                    &sleaf_, &spre_, &kdoff_, &rtop_, &lines_, &tickets_, &abort_, &xl_off_,
                    &xl_total_, &xl_words_, &tail_lv_})
    b->release();
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : prof_ev_) (void)hipEventDestroy(e);
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  for (auto& e : join_ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& sd : side_)
    if (sd) (void)hipStreamDestroy(sd);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int Context::init() {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    set_error("no HIP device available");
    return ERR_DEVICE;
  }
  if (device_ < 0 || device_ >= count) {
    set_error("device index out of range");
    return ERR_INVALID_ARG;
  }
  HIP_TRY(hipSetDevice(device_));
  HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  for (auto& sd : side_) HIP_TRY(hipStreamCreateWithFlags(&sd, hipStreamNonBlocking));
  for (auto& ev : ev_) HIP_TRY(hipEventCreate(&ev));
  HIP_TRY(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
  for (auto& ev : join_ev_) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  // Slot storage limit: what the device grants one workgroup, less the
  // row-scan kernels' static LDS, capped at kLdsMaxTable.
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_));
  lds_limit = (int32_t)std::min<size_t>(kLdsMaxTable, prop.sharedMemPerBlock - 256);
  HIP_TRY(forest_init_attributes(lds_limit));
  rule.lds_max = lds_limit;
  return OK;
}

PoolView Context::view() const {
  PoolView v;
  v.parent = parent_.as<int32_t>();
  v.label = label_.as<int32_t>();
  v.orl = orl_.as<int32_t>();
  v.kdep = kdep_.as<int32_t>();
  v.kdn = kdn_.as<int32_t>();
  v.lrank = lrank_.as<int32_t>();
  v.kr = kr_.as<int32_t>();
  v.krkd = krkd_.as<int32_t>();
  v.krsize = krsize_.as<int32_t>();
  v.krleaf = krleaf_.as<int32_t>();
  v.skr = skr_.as<int32_t>();
  v.skd = skd_.as<int32_t>();
  v.ssize = ssize_.as<int32_t>();
  v.sleaf = sleaf_.as<int32_t>();
  v.spre = spre_.as<int32_t>();
  v.kdoff = kdoff_.as<int32_t>();
  v.rtop = rtop_.as<int32_t>();
  v.trees = trees_.as<TreeMeta>();
  return v;
}

int Context::grow_pool(int64_t nodes) {
  const size_t bytes = (size_t)nodes * sizeof(int32_t);
  for (DevBuf* b : {&parent_, &label_, &orl_, &kdep_, &kdn_, &lrank_, &kr_, &krkd_, &krsize_,
                    &krleaf_, &skr_, &skd_, &ssize_, &sleaf_, &spre_, &kdoff_, &rtop_})
    ST_TRY(b->grow(bytes, true, stream_));
  return OK;
}

int Context::pool_clear() {
  std::lock_guard<std::mutex> lk(mu_);
  meta_.clear();
  info_.clear();
  pool_nodes_ = 0;
  return OK;
}

int Context::pool_add(int32_t ntrees, const int32_t* sizes, const int32_t* parents,
                      const int32_t* labels, int32_t* ids) {
  std::lock_guard<std::mutex> lk(mu_);
  if (ntrees < 0 || (ntrees > 0 && (!sizes || !parents || !labels))) {
    set_error("pool_add: null argument");
    return ERR_INVALID_ARG;
  }
  if (ntrees == 0) return OK;
  HIP_TRY(hipSetDevice(device_));
  // Validate on the host and lay out the padded staging arrays.
  int64_t total = 0, in_off = 0, scratch_total = 0;
  std::vector<int64_t> offs(ntrees), scr(ntrees);
  for (int32_t t = 0; t < ntrees; ++t) {
    const int32_t n = sizes[t];
    if (n <= 0) {
      set_error("pool_add: tree " + std::to_string(t) + " is empty");
      return ERR_INVALID_ARG;
    }
    if (n > kMaxTreeNodes) {
      set_error("pool_add: tree " + std::to_string(t) + " has " + std::to_string(n) +
                " nodes; the limit is " + std::to_string(kMaxTreeNodes));
      return ERR_TOO_LARGE;
    }
    if (validate_preorder(parents + in_off, n) != OK) {
      set_error("pool_add: tree " + std::to_string(t) + " is not a preorder parent array");
      return ERR_NOT_PREORDER;
    }
    offs[t] = pool_nodes_ + total;
    scr[t] = scratch_total;
    total += n + 1;
    scratch_total += 6 * (int64_t)(n + 1);
    in_off += n;
  }
  std::vector<int32_t> hp(total, -1), hl(total, 0);
  std::vector<TreeMeta> tm(ntrees);
  std::vector<HostTreeInfo> ti(ntrees);
  in_off = 0;
  for (int32_t t = 0; t < ntrees; ++t) {
    const int32_t n = sizes[t];
    const int64_t o = offs[t] - pool_nodes_;
    std::memcpy(&hp[o], parents + in_off, sizeof(int32_t) * n);
    std::memcpy(&hl[o], labels + in_off, sizeof(int32_t) * n);
    tm[t] = TreeMeta{offs[t], n, 0, 0, 0, 0, 0};
    // Host-side cell count for reporting only.
    std::vector<int32_t> orl(n);
    compute_orl(parents + in_off, n, orl.data());
    ti[t].n = n;
    ti[t].ksum = keyroot_size_sum(orl.data(), keyroots_desc(parents + in_off, orl.data(), n));
    in_off += n;
  }
  const int32_t first = (int32_t)meta_.size();
  ST_TRY(grow_pool(pool_nodes_ + total));
  ST_TRY(trees_.grow(sizeof(TreeMeta) * (first + ntrees), true, stream_));
  ST_TRY(prep_scratch_.grow(sizeof(int32_t) * scratch_total, false, stream_));
  ST_TRY(prep_off_.grow(sizeof(int64_t) * ntrees, false, stream_));
  HIP_TRY(hipMemcpyAsync(parent_.as<int32_t>() + pool_nodes_, hp.data(), sizeof(int32_t) * total,
                         hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipMemcpyAsync(label_.as<int32_t>() + pool_nodes_, hl.data(), sizeof(int32_t) * total,
                         hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipMemcpyAsync(trees_.as<TreeMeta>() + first, tm.data(), sizeof(TreeMeta) * ntrees,
                         hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipMemcpyAsync(prep_off_.p, scr.data(), sizeof(int64_t) * ntrees, hipMemcpyHostToDevice,
                         stream_));
  PrepArgs pa;
  pa.parent = parent_.as<int32_t>();
  pa.label = label_.as<int32_t>();
  pa.orl = orl_.as<int32_t>();
  pa.kdep = kdep_.as<int32_t>();
  pa.kdn = kdn_.as<int32_t>();
  pa.lrank = lrank_.as<int32_t>();
  pa.kr = kr_.as<int32_t>();
  pa.krkd = krkd_.as<int32_t>();
  pa.krsize = krsize_.as<int32_t>();
  pa.krleaf = krleaf_.as<int32_t>();
  pa.skr = skr_.as<int32_t>();
  pa.skd = skd_.as<int32_t>();
  pa.ssize = ssize_.as<int32_t>();
  pa.sleaf = sleaf_.as<int32_t>();
  pa.spre = spre_.as<int32_t>();
  pa.kdoff = kdoff_.as<int32_t>();
  pa.rtop = rtop_.as<int32_t>();
  pa.trees = trees_.as<TreeMeta>();
  pa.scratch = prep_scratch_.as<int32_t>();
  pa.scratch_off = prep_off_.as<int64_t>();
  pa.first_tree = first;
  launch_tree_prep(pa, ntrees, stream_);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(tm.data(), trees_.as<TreeMeta>() + first, sizeof(TreeMeta) * ntrees,
                         hipMemcpyDeviceToHost, stream_));
  HIP_TRY(hipStreamSynchronize(stream_));
  for (int32_t t = 0; t < ntrees; ++t) {
    meta_.push_back(tm[t]);
    info_.push_back(ti[t]);
    if (ids) ids[t] = first + t;
  }
  pool_nodes_ += total;
  return OK;
}

int Context::tree_arrays(int32_t id, int32_t* orl, int32_t* kr, int32_t* K, int32_t* kd) {
  std::lock_guard<std::mutex> lk(mu_);
  if (id < 0 || id >= (int32_t)meta_.size()) {
    set_error("tree id out of range");
    return ERR_NOT_FOUND;
  }
  HIP_TRY(hipSetDevice(device_));
  const TreeMeta& m = meta_[id];
  if (orl)
    HIP_TRY(hipMemcpyAsync(orl, orl_.as<int32_t>() + m.node_off, sizeof(int32_t) * m.n,
                           hipMemcpyDeviceToHost, stream_));
  if (kr)
    HIP_TRY(hipMemcpyAsync(kr, kr_.as<int32_t>() + m.node_off, sizeof(int32_t) * m.K,
                           hipMemcpyDeviceToHost, stream_));
  if (kd)
    HIP_TRY(hipMemcpyAsync(kd, krkd_.as<int32_t>() + m.node_off, sizeof(int32_t) * m.K,
                           hipMemcpyDeviceToHost, stream_));
  HIP_TRY(hipStreamSynchronize(stream_));
  if (K) *K = m.K;
  return OK;
}

int Context::run_batch(int32_t npairs, const int32_t* xs, const int32_t* ys, int32_t* dist,
                       int32_t* const* td_out, BatchStats* stats) {
  std::lock_guard<std::mutex> lk(mu_);
  const auto t_start = std::chrono::steady_clock::now();
  BatchStats st;
  if (npairs < 0 || (npairs > 0 && (!xs || !ys || !dist))) {
    set_error("run_batch: null argument");
    return ERR_INVALID_ARG;
  }
  if (npairs == 0) {
    if (stats) *stats = st;
    return OK;
  }
  if (npairs >= (1 << kTaskPairBits)) {
    set_error("run_batch: too many pairs in one batch");
    return ERR_TOO_LARGE;
  }
  HIP_TRY(hipSetDevice(device_));
  std::vector<PairMeta> pm(npairs);
  std::vector<int64_t> blk(npairs + 1, 0), lines(npairs + 1, 0);
  int64_t td_total = 0, planned = 0, leaf_tables = 0, leaf_cells = 0;
  int32_t nlevels = 0, maxn = 1;
  for (int32_t i = 0; i < npairs; ++i) {
    const int32_t x = xs[i], y = ys[i];
    if (x < 0 || y < 0 || x >= (int32_t)meta_.size() || y >= (int32_t)meta_.size()) {
      set_error("run_batch: tree id out of range in pair " + std::to_string(i));
      return ERR_NOT_FOUND;
    }
    const TreeMeta &a = meta_[x], &b = meta_[y];
    const int64_t mn = (int64_t)a.n * b.n;
    maxn = std::max(maxn, std::max(a.n, b.n));
    pm[i] = PairMeta{x, y, td_total + kTdPad, td_total + 2 * kTdPad + mn};
    td_total += 2 * mn + 3 * kTdPad;
    // With the leaf kernel on, the plan covers only non-leaf keyroot pairs;
    // the leaf kernel writes one line per leaf keyroot of either tree.
    const int32_t ka = lane_enable ? a.K - a.K1 : a.K, kb = lane_enable ? b.K - b.K1 : b.K;
    lines[i + 1] = lines[i] + (lane_enable ? a.K1 + b.K1 : 0);
    if (lane_enable) {
      leaf_tables += (int64_t)a.K1 * b.K + (int64_t)(a.K - a.K1) * b.K1;
      leaf_cells += (int64_t)a.K1 * info_[y].ksum + (int64_t)b.K1 * (info_[x].ksum - a.K1);
    }
    blk[i + 1] = blk[i] + plan_blocks(ka, kb, plan_chunk);
    planned += (int64_t)ka * kb;
    nlevels = std::max(nlevels, a.kdmax + b.kdmax + 1);
    st.tables += (int64_t)a.K * b.K;
    st.cells += info_[x].ksum * info_[y].ksum;
  }
  const int nb = nlevels * kNumClasses;
  // Levels that at most tail_pairs pairs reach use the latency rule.
  std::vector<int32_t> active(nlevels + 1, 0);
  for (int32_t i = 0; i < npairs; ++i) ++active[meta_[xs[i]].kdmax + meta_[ys[i]].kdmax];
  h_tail_.assign(nlevels, 0);
  for (int L = nlevels - 1, run = 0; L >= 0; --L) {
    run += active[L];
    h_tail_[L] = (tail_pairs > 0 && run <= tail_pairs) ? 1 : 0;
  }
  ST_TRY(pairs_.grow(sizeof(PairMeta) * npairs, false, stream_));
  ST_TRY(tail_lv_.grow(nlevels, false, stream_));
  ST_TRY(blk_.grow(sizeof(int64_t) * (npairs + 1), false, stream_));
  ST_TRY(counts_.grow(sizeof(unsigned long long) * nb, false, stream_));
  ST_TRY(cells_.grow(sizeof(unsigned long long) * nb, false, stream_));
  ST_TRY(cursor_.grow(sizeof(unsigned long long) * nb, false, stream_));
  ST_TRY(maxlds_.grow(sizeof(int32_t) * nb, false, stream_));
  ST_TRY(tickets_.grow(sizeof(unsigned int) * nb, false, stream_));
  ST_TRY(abort_.grow(sizeof(int32_t), false, stream_));
  ST_TRY(tasks_.grow(sizeof(uint64_t) * std::max<int64_t>(planned, 1), false, stream_));
  ST_TRY(td_.grow(sizeof(int32_t) * td_total, false, stream_));
  ST_TRY(dist_.grow(sizeof(int32_t) * npairs, false, stream_));
  ST_TRY(lines_.grow(sizeof(int64_t) * (npairs + 1), false, stream_));
  HIP_TRY(hipMemcpyAsync(pairs_.p, pm.data(), sizeof(PairMeta) * npairs, hipMemcpyHostToDevice,
                         stream_));
  HIP_TRY(hipMemcpyAsync(blk_.p, blk.data(), sizeof(int64_t) * (npairs + 1),
                         hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipMemcpyAsync(lines_.p, lines.data(), sizeof(int64_t) * (npairs + 1),
                         hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipMemcpyAsync(tail_lv_.p, h_tail_.data(), nlevels, hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipEventRecord(ev_[0], stream_));

  PlanArgs pa;
  pa.pool = view();
  pa.pairs = pairs_.as<PairMeta>();
  pa.blk_start = blk_.as<int64_t>();
  pa.npairs = npairs;
  pa.nlevels = nlevels;
  pa.lchunk_tables = plan_chunk;
  pa.rule = rule;
  pa.rule_tail = ClassRule{lds_limit, 1, 1 << 30, lds_limit, 16};
  pa.tail_level = tail_lv_.as<uint8_t>();
  pa.force_class = force_class;
  pa.lane_enable = lane_enable;
  pa.counts = counts_.as<unsigned long long>();
  pa.cells = cells_.as<unsigned long long>();
  pa.cursor = cursor_.as<unsigned long long>();
  pa.maxlds = maxlds_.as<int32_t>();
  pa.tasks = tasks_.as<uint64_t>();
  pa.tickets = tickets_.as<unsigned int>();
  pa.abort_flag = abort_.as<int32_t>();
  launch_plan(pa, blk[npairs], stream_);
  HIP_TRY(hipGetLastError());
  LeafArgs la;
  la.pool = pa.pool;
  la.pairs = pa.pairs;
  la.line_start = lines_.as<int64_t>();
  la.npairs = npairs;
  la.td = td_.as<int32_t>();
  h_counts_.assign(nb, 0);
  h_maxlds_.assign(nb, 0);
  HIP_TRY(hipMemcpyAsync(h_counts_.data(), counts_.p, sizeof(unsigned long long) * nb,
                         hipMemcpyDeviceToHost, stream_));
  HIP_TRY(hipMemcpyAsync(h_maxlds_.data(), maxlds_.p, sizeof(int32_t) * nb,
                         hipMemcpyDeviceToHost, stream_));
  h_cells_.assign(nb, 0);
  HIP_TRY(hipMemcpyAsync(h_cells_.data(), cells_.p, sizeof(unsigned long long) * nb,
                         hipMemcpyDeviceToHost, stream_));
  HIP_TRY(hipStreamSynchronize(stream_));
  HIP_TRY(hipEventRecord(ev_[1], stream_));
  // Column-blocked tables: ordinal of each bucket's first task, then (with a
  // second short sync, only in batches that have such tables) the arena
  // layout computed on the device.
  std::vector<int64_t> xl_first(nb, 0);
  int64_t nxl = 0;
  for (int L = 0; L < nlevels; ++L) {
    if (h_counts_[(size_t)L * kNumClasses + kClassXL]) {
      set_error("run_batch: a table fits no form (nesting height above " +
                std::to_string(kMaxImports - 1) + " or slot storage above " +
                std::to_string(lds_limit) + " bytes even in 256-column blocks)");
      return ERR_TOO_LARGE;
    }
    for (int k = 0; k < kNumXLForms; ++k) {
      const size_t bkt = (size_t)L * kNumClasses + kFirstXLClass + k;
      xl_first[bkt] = nxl;
      nxl += (int64_t)h_counts_[bkt];
    }
  }
  if (nxl > 0) {
    ST_TRY(xl_off_.grow(sizeof(int64_t) * nxl, false, stream_));
    ST_TRY(xl_total_.grow(sizeof(unsigned long long), false, stream_));
    XLLayoutArgs xa;
    xa.pool = pa.pool;
    xa.pairs = pa.pairs;
    xa.tasks = pa.tasks;
    xa.counts = pa.counts;
    xa.cursor = pa.cursor;
    xa.nlevels = nlevels;
    xa.off = xl_off_.as<int64_t>();
    xa.total = xl_total_.as<unsigned long long>();
    launch_xl_layout(xa, stream_);
    HIP_TRY(hipGetLastError());
    unsigned long long words = 0;
    HIP_TRY(hipMemcpyAsync(&words, xa.total, sizeof(words), hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    const size_t bytes = sizeof(unsigned long long) * std::max<unsigned long long>(words, 1);
    if (bytes > xl_words_.cap) {
      ST_TRY(xl_words_.grow(bytes, false, stream_));
      launch_xl_zero(xl_words_.as<unsigned long long>(), (int64_t)(xl_words_.cap / 8), stream_);
      HIP_TRY(hipGetLastError());
    }
  }
  DPArgs da;
  da.pool = pa.pool;
  da.pairs = pa.pairs;
  da.td = td_.as<int32_t>();
  da.xl_words = xl_words_.as<unsigned long long>();
  da.xl_abort = abort_.as<int32_t>();
  // With profiling on, an event pair brackets every launch (same stream).
  std::vector<std::pair<int, int>> prof;  // (class, bucket or -1-level) per timed launch
  if (profile) {
    size_t need = 0;
    for (unsigned long long c : h_counts_) need += c ? 2 : 0;
    need += 2;
    while (prof_ev_.size() < need) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      prof_ev_.push_back(e);
    }
  }
  // Launches of one level are independent: they fork from the main stream
  // onto up to level_streams streams (largest first) and join before the
  // next level, so latency-bound launches of few large tables overlap with
  // the rest of their level instead of queueing behind it.
  struct Launch {
    int cls, bucket;
    unsigned long long off, cnt, cells;
  };
  std::vector<Launch> lv;
  unsigned long long off = 0;
  for (int L = 0; L < nlevels; ++L) {
    lv.clear();
    // The leaf lines read nothing and level-0 tables (two chains) read no TD,
    // so the leaf kernel runs beside level 0.
    if (L == 0 && lines[npairs] > 0)
      lv.push_back({kClassLane, -1, 0, (unsigned long long)lines[npairs],
                    (unsigned long long)leaf_cells});
    for (int c = 0; c < kNumClasses; ++c) {
      const size_t b = (size_t)L * kNumClasses + c;
      if (h_counts_[b]) lv.push_back({c, (int)b, off, h_counts_[b], h_cells_[b]});
      off += h_counts_[b];
    }
    std::stable_sort(lv.begin(), lv.end(),
                     [](const Launch& x, const Launch& y) { return x.cells > y.cells; });
    const int ns = std::min<int>((int)lv.size(), std::max(1, level_streams));
    if (ns > 1) HIP_TRY(hipEventRecord(fork_ev_, stream_));
    for (int k = 1; k < ns; ++k) HIP_TRY(hipStreamWaitEvent(side_[k - 1], fork_ev_, 0));
    for (size_t i = 0; i < lv.size(); ++i) {
      const Launch& w = lv[i];
      hipStream_t s = (i % ns) == 0 ? stream_ : side_[i % ns - 1];
      if (profile) HIP_TRY(hipEventRecord(prof_ev_[2 * prof.size()], s));
      if (w.cls == kClassLane) {
        launch_leaf(la, (int64_t)w.cnt, s);
      } else {
        da.tasks = tasks_.as<uint64_t>() + w.off;
        da.ntasks = (int64_t)w.cnt;
        if (w.cls >= kFirstXLClass) {
          const int bw = xl_bw(w.cls - kFirstXLClass);
          da.xl_off = xl_off_.as<int64_t>() + xl_first[w.bucket];
          da.xl_ticket = tickets_.as<unsigned int>() + w.bucket;
          da.xl_epoch = ++xl_epoch_ ? xl_epoch_ : ++xl_epoch_;  // never 0
          da.xl_nbmax = (maxn + bw - 1) / bw;
        }
        const int lds = std::max(64, (h_maxlds_[w.bucket] + 15) & ~15);
        if (!launch_forest(w.cls, da, lds, s)) {
          set_error("run_batch: no kernel for class " + std::to_string(w.cls));
          return ERR_INVALID_ARG;
        }
      }
      if (profile) {
        HIP_TRY(hipEventRecord(prof_ev_[2 * prof.size() + 1], s));
        prof.emplace_back(w.cls, w.bucket);
      }
      ++st.launches;
    }
    for (int k = 1; k < ns; ++k) {
      HIP_TRY(hipEventRecord(join_ev_[k - 1], side_[k - 1]));
      HIP_TRY(hipStreamWaitEvent(stream_, join_ev_[k - 1], 0));
    }
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ev_[2], stream_));
  launch_gather_dist(pa.pairs, da.td, npairs, dist_.as<int32_t>(), stream_);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(dist, dist_.p, sizeof(int32_t) * npairs, hipMemcpyDeviceToHost, stream_));
  int32_t aborted = 0;
  if (nxl > 0)
    HIP_TRY(hipMemcpyAsync(&aborted, abort_.p, sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
  if (td_out) {
    for (int32_t i = 0; i < npairs; ++i) {
      if (!td_out[i]) continue;
      const int64_t mn = (int64_t)meta_[xs[i]].n * meta_[ys[i]].n;
      HIP_TRY(hipMemcpyAsync(td_out[i], da.td + pm[i].td_off, sizeof(int32_t) * mn,
                             hipMemcpyDeviceToHost, stream_));
    }
  }
  HIP_TRY(hipStreamSynchronize(stream_));
  if (aborted) {
    set_error("run_batch: a column-blocked table stopped waiting for its neighbour block");
    return ERR_TIMEOUT;
  }
  if (profile) {
    for (auto& cp : class_profile) cp = ClassProfile{};
    for (size_t i = 0; i < prof.size(); ++i) {
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, prof_ev_[2 * i], prof_ev_[2 * i + 1]));
      ClassProfile& cp = class_profile[prof[i].first];
      cp.launches += 1;
      cp.ms += ms;
      const int b = prof[i].second;
      cp.tables += b >= 0 ? (int64_t)h_counts_[b] : leaf_tables;
      cp.cells += b >= 0 ? (int64_t)h_cells_[b] : leaf_cells;
    }
  }
  float plan_ms = 0, dp_ms = 0;
  HIP_TRY(hipEventElapsedTime(&plan_ms, ev_[0], ev_[1]));
  HIP_TRY(hipEventElapsedTime(&dp_ms, ev_[1], ev_[2]));
  st.plan_ms = plan_ms;
  st.dp_ms = dp_ms;
  st.levels = nlevels;
  st.total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (stats) *stats = st;
  return OK;
}

}  // namespace xted
