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
#include <array>
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
  if (arena) {
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (arena->used + need > arena->cap) {
      set_error("no-allocation batch: the pairs need more than the " +
                std::to_string(arena->cap) + " bytes of the device scratch array (" +
                std::to_string(arena->used + need) + " bytes so far)");
      return ERR_TOO_LARGE;
    }
    p = arena->base + arena->used;
    arena->used += need;
    cap = need;
    return OK;
  }
  size_t ncap = std::max(bytes, cap + cap / 2);
  ncap = (ncap + 4095) & ~(size_t)4095;
  void* np = nullptr;
  hipError_t e = fine ? hipExtMallocWithFlags(&np, ncap, hipDeviceMallocFinegrained)
                      : hipMalloc(&np, ncap);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("device allocation of " + std::to_string(ncap) + " bytes failed: " +
              hipGetErrorString(e));
    return ERR_OUT_OF_MEMORY;
  }
  if (p) {
    if (keep && cap) HIP_TRY(hipMemcpyAsync(np, p, cap, hipMemcpyDeviceToDevice, s));
    if (retire && !keep) {
      retire->push_back(p);
    } else {
      HIP_TRY(hipStreamSynchronize(s));
      HIP_TRY(hipFree(p));
    }
  }
  p = np;
  cap = ncap;
  return OK;
}

void DevBuf::release() {
  if (p && !arena) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

// The no-allocation batch path's buffers, one array per device (each device
// holds its own copy). Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (Context::init).
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

Context::Context(int device) : device_(device) {}

Context::~Context() {
  if (stream_) {
    (void)hipSetDevice(device_);
    (void)hipDeviceSynchronize();
  }
  for (Group& g : batch_) g.release();
  na_.release();
  for (DevBuf* b : {&parent_, &label_, &orl_, &kdep_, &kdn_, &lrank_, &kr_, &krkd_, &krsize_,
                    &krleaf_, &trees_, &prep_scratch_, &prep_off_, &skr_, &skd_, &ssize_, &sleaf_,
                    &spre_, &kdoff_, &rtop_, &cdel_, &cins_, &cren_, &pdel_, &pins_, &ctot_, &crec_,
                    &rrec_, &ccrec_})
    b->release();
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : prof_ev_) (void)hipEventDestroy(e);
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  if (plan_done_) (void)hipEventDestroy(plan_done_);
  for (auto& e : join_ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& sd : side_)
    if (sd) (void)hipStreamDestroy(sd);
  if (plan_stream_) (void)hipStreamDestroy(plan_stream_);
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
  HIP_TRY(hipStreamCreateWithFlags(&plan_stream_, hipStreamNonBlocking));
  for (auto& sd : side_) HIP_TRY(hipStreamCreateWithFlags(&sd, hipStreamNonBlocking));
  for (auto& ev : ev_) HIP_TRY(hipEventCreate(&ev));
  HIP_TRY(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&plan_done_, hipEventDisableTiming));
  for (auto& ev : join_ev_) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  // Slot storage limit: what the device grants one workgroup, less the
  // row-scan kernels' static LDS, capped at kLdsMaxTable.
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_));
  lds_limit = (int32_t)std::min<size_t>(kLdsMaxTable, prop.sharedMemPerBlock - 256);
  HIP_TRY(forest_init_attributes(lds_limit));
  rule.lds_max = lds_limit;
  // This device's copy of the no-allocation array, and what its batch group
  // needs besides memory, so that run_batch_noalloc creates nothing.
  void* sp = nullptr;
  HIP_TRY(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  na_arena_ = Arena{static_cast<char*>(sp), kNoAllocBytes, 0};
  na_.for_each_buf([this](DevBuf& b) { b.arena = &na_arena_; });
  for (hipEvent_t* e : {&na_.ev_plan0, &na_.ev_plan1, &na_.ev_end}) HIP_TRY(hipEventCreate(e));
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
  v.pdel = pdel_.as<int32_t>();
  v.pins = pins_.as<int32_t>();
  v.crec = crec_.as<NodeRec>();
  v.rrec = rrec_.as<NodeRec>();
  v.ccrec = ccrec_.as<CostRec>();
  v.trees = trees_.as<TreeMeta>();
  return v;
}

int Context::grow_pool(int64_t nodes) {
  const size_t bytes = (size_t)nodes * sizeof(int32_t);
  for (DevBuf* b : {&parent_, &label_, &orl_, &kdep_, &kdn_, &lrank_, &kr_, &krkd_, &krsize_,
                    &krleaf_, &skr_, &skd_, &ssize_, &sleaf_, &spre_, &kdoff_, &rtop_, &pdel_,
                    &pins_})
    ST_TRY(b->grow(bytes, true, stream_));
  for (DevBuf* b : {&crec_, &rrec_})
    ST_TRY(b->grow((size_t)nodes * sizeof(NodeRec), true, stream_));
  ST_TRY(ccrec_.grow((size_t)nodes * sizeof(CostRec), true, stream_));
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
    ti[t].line_stack = line_stack_depth(parents + in_off, orl.data(), n);
    ti[t].maxlab = *std::max_element(labels + in_off, labels + in_off + n);
    ti[t].minlab = *std::min_element(labels + in_off, labels + in_off + n);
    if (cost_alphabet_ > 0 && (ti[t].minlab < 0 || ti[t].maxlab >= cost_alphabet_)) {
      set_error("pool_add: tree " + std::to_string(t) + " has a label outside the cost alphabet [0, " +
                std::to_string(cost_alphabet_) + ")");
      return ERR_INVALID_ARG;
    }
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
  pa.crec = crec_.as<NodeRec>();
  pa.rrec = rrec_.as<NodeRec>();
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
  if (cost_alphabet_ > 0) return cost_prefix(first, ntrees);
  return OK;
}

int Context::cost_prefix(int32_t first, int32_t ntrees) {
  if (ntrees <= 0) return OK;
  ST_TRY(ctot_.grow(sizeof(int32_t) * 2 * ntrees, false, stream_));
  CostPrefixArgs ca;
  ca.trees = trees_.as<TreeMeta>();
  ca.label = label_.as<int32_t>();
  ca.del = cdel_.as<int32_t>();
  ca.ins = cins_.as<int32_t>();
  ca.pdel = pdel_.as<int32_t>();
  ca.pins = pins_.as<int32_t>();
  ca.totals = ctot_.as<int32_t>();
  ca.orl = orl_.as<int32_t>();
  ca.rec = ccrec_.as<CostRec>();
  ca.first_tree = first;
  ca.alphabet = cost_alphabet_;
  launch_cost_prefix(ca, ntrees, stream_);
  HIP_TRY(hipGetLastError());
  std::vector<int32_t> tot(2 * ntrees);
  HIP_TRY(hipMemcpyAsync(tot.data(), ctot_.p, sizeof(int32_t) * 2 * ntrees, hipMemcpyDeviceToHost,
                         stream_));
  HIP_TRY(hipStreamSynchronize(stream_));
  for (int32_t t = 0; t < ntrees; ++t) {
    info_[first + t].sdel = tot[2 * t];
    info_[first + t].sins = tot[2 * t + 1];
  }
  return OK;
}

int Context::set_costs(int32_t alphabet, const int32_t* del, const int32_t* ins,
                       const int32_t* ren) {
  std::lock_guard<std::mutex> lk(mu_);
  // Work already launched finishes under the old model before its buffers
  // change; groups planned under it and not yet fully launched are refused
  // at their next launch.
  HIP_TRY(hipSetDevice(device_));
  HIP_TRY(hipDeviceSynchronize());
  if (!del) {
    ++cost_epoch_;
    cost_alphabet_ = 0;
    return OK;
  }
  if (alphabet < 1 || alphabet > 4096 || !ins || !ren) {
    set_error("set_costs: alphabet must be 1 to 4096 with del, ins and ren given");
    return ERR_INVALID_ARG;
  }
  const int64_t a2 = (int64_t)alphabet * alphabet;
  auto bad = [](const int32_t* p, int64_t k) {
    for (int64_t i = 0; i < k; ++i)
      if (p[i] < 0 || p[i] > 32767) return true;
    return false;
  };
  if (bad(del, alphabet) || bad(ins, alphabet) || bad(ren, a2)) {
    set_error("set_costs: costs must lie in [0, 32767]");
    return ERR_INVALID_ARG;
  }
  ++cost_epoch_;
  cost_alphabet_ = 0;  // until the new model is in place
  ST_TRY(cdel_.grow(sizeof(int32_t) * alphabet, false, stream_));
  ST_TRY(cins_.grow(sizeof(int32_t) * alphabet, false, stream_));
  ST_TRY(cren_.grow(sizeof(int32_t) * a2, false, stream_));
  HIP_TRY(hipMemcpyAsync(cdel_.p, del, sizeof(int32_t) * alphabet, hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipMemcpyAsync(cins_.p, ins, sizeof(int32_t) * alphabet, hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipMemcpyAsync(cren_.p, ren, sizeof(int32_t) * a2, hipMemcpyHostToDevice, stream_));
  HIP_TRY(hipStreamSynchronize(stream_));
  hdel_.assign(del, del + alphabet);
  hins_.assign(ins, ins + alphabet);
  cost_alphabet_ = alphabet;
  return cost_prefix(0, (int32_t)meta_.size());
}

int64_t Context::whole_tree_cost(const int32_t* labels, int32_t n, bool del) const {
  int64_t s = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (labels[i] < 0 || labels[i] >= cost_alphabet_) return -1;
    s += del ? hdel_[labels[i]] : hins_[labels[i]];
  }
  return s;
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

int xl_timeout(const int32_t* rec);

void Group::release() {
  if (h_dist_pin) (void)hipHostFree(h_dist_pin);
  h_dist_pin = nullptr;
  pin_cap = 0;
  for (hipEvent_t* e : {&done_ev, &ev_plan0, &ev_plan1, &ev_end}) {
    if (*e) (void)hipEventDestroy(*e);
    *e = nullptr;
  }
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  graph_exec = nullptr;
  graph = nullptr;
  graph_sig.clear();
  for_each_buf([](DevBuf& b) { b.release(); });
}

int Context::lane_streams(int lane, hipStream_t* out) const {
  if (lane >= 0 && lane < kMaxLevelStreams) {
    out[0] = lane == 0 ? stream_ : side_[lane - 1];
    return 1;
  }
  out[0] = stream_;
  for (int k = 1; k < kMaxLevelStreams; ++k) out[k] = side_[k - 1];
  return std::max(1, std::min(level_streams, kMaxLevelStreams));
}

hipStream_t Context::lane_stream(int lane) const {
  hipStream_t s[kMaxLevelStreams];
  lane_streams(lane, s);
  return s[0];
}

int Context::plan(Group& g, int32_t npairs, const int32_t* xs, const int32_t* ys, int lane) {
  BatchStats& st = g.st;
  st = BatchStats{};
  g.npairs = 0;
  if (npairs <= 0 || !xs || !ys) {
    set_error("plan: null argument");
    return ERR_INVALID_ARG;
  }
  if (npairs >= (1 << kTaskPairBits)) {
    set_error("run_batch: too many pairs in one batch");
    return ERR_TOO_LARGE;
  }
  HIP_TRY(hipSetDevice(device_));
  // The plan runs on its own stream, so a server can plan a new group while
  // the levels of others run; it returns once the plan is on the host.
  hipStream_t ps = plan_stream_;
  g.xs.assign(xs, xs + npairs);
  g.ys.assign(ys, ys + npairs);
  g.pm.resize(npairs);
  std::vector<int64_t> blk(npairs + 1, 0);
  g.lines.assign(npairs + 1, 0);
  int64_t td_total = 0, planned = 0;
  g.leaf_tables = g.leaf_cells = 0;
  int32_t nlevels = 0;
  g.maxn = 1;
  // Pairs whose leaf lines are in closed form (per pair under a cost model,
  // whose form needs both trees within its stack); the others plan the tables
  // of their leaf keyroots like the rest.
  std::vector<uint8_t> leaf(npairs, 0);
  for (int32_t i = 0; i < npairs; ++i) {
    const int32_t x = xs[i], y = ys[i];
    if (x < 0 || y < 0 || x >= (int32_t)meta_.size() || y >= (int32_t)meta_.size()) {
      set_error("run_batch: tree id out of range in pair " + std::to_string(i));
      return ERR_NOT_FOUND;
    }
    const bool le = leaf_form(x, y);
    leaf[i] = le ? 1 : 0;
    const TreeMeta &a = meta_[x], &b = meta_[y];
    const int64_t mn = (int64_t)a.n * b.n;
    // Far values sit in 16-bit LDS slots: every forest value (at most
    // deleting all of x and inserting all of y) must stay below 2^16.
    if (cost_alphabet_ > 0 && (info_[x].minlab < 0 || info_[x].maxlab >= cost_alphabet_ ||
                               info_[y].minlab < 0 || info_[y].maxlab >= cost_alphabet_)) {
      set_error("run_batch: pair " + std::to_string(i) +
                " uses a pool tree with labels outside the cost alphabet");
      return ERR_INVALID_ARG;
    }
    if (cost_alphabet_ > 0 && info_[x].sdel + info_[y].sins > 65535) {
      set_error("run_batch: pair " + std::to_string(i) + " can cost " +
                std::to_string(info_[x].sdel + info_[y].sins) +
                " under the cost model; forest values above 65535 are not supported");
      return ERR_TOO_LARGE;
    }
    g.maxn = std::max(g.maxn, std::max(a.n, b.n));
    g.pm[i] = PairMeta{x, y, td_total + kTdPad, td_total + 2 * kTdPad + mn};
    td_total += 2 * mn + 3 * kTdPad;
    // With the leaf kernel on, the plan covers only non-leaf keyroot pairs;
    // the leaf kernel writes one line per leaf keyroot of either tree, in
    // tiles of up to 64 lines of one side.
    const int32_t ka = le ? a.K - a.K1 : a.K, kb = le ? b.K - b.K1 : b.K;
    g.lines[i + 1] = g.lines[i] + (le ? (a.n + 63) / 64 + (b.n + 63) / 64 : 0);
    if (le) {
      g.leaf_tables += (int64_t)a.K1 * b.K + (int64_t)(a.K - a.K1) * b.K1;
      g.leaf_cells += (int64_t)a.K1 * info_[y].ksum + (int64_t)b.K1 * (info_[x].ksum - a.K1);
    }
    blk[i + 1] = blk[i] + plan_blocks(ka, kb, plan_chunk);
    planned += (int64_t)ka * kb;
    nlevels = std::max(nlevels, a.kdmax + b.kdmax + 1);
    st.tables += (int64_t)a.K * b.K;
    st.cells += info_[x].ksum * info_[y].ksum;
  }
  g.td_bytes = (int64_t)sizeof(td_t) * td_total;
  const int nb = nlevels * kNumClasses * kBands;
  g.loff.assign(npairs, 0);
  if (level_align == 1 || level_align == 2) {
    for (int32_t i = 0; i < npairs; ++i) {
      const int32_t spare = nlevels - (meta_[xs[i]].kdmax + meta_[ys[i]].kdmax + 1);
      g.loff[i] = level_align == 1 ? spare : spare / 2;
    }
  }
  // Levels that at most tail_pairs pairs reach use the latency rule.
  std::vector<int32_t> active(nlevels + 1, 0);
  for (int32_t i = 0; i < npairs; ++i)
    ++active[meta_[xs[i]].kdmax + meta_[ys[i]].kdmax + g.loff[i]];
  g.h_tail.assign(nlevels, 0);
  for (int L = nlevels - 1, run = 0; L >= 0; --L) {
    run += active[L];
    g.h_tail[L] = (tail_pairs > 0 && (int64_t)run * plan_share <= tail_pairs) ? 1 : 0;
  }
  ST_TRY(g.pairs.grow(sizeof(PairMeta) * npairs, false, ps));
  ST_TRY(g.tail_lv.grow(nlevels, false, ps));
  ST_TRY(g.blk.grow(sizeof(int64_t) * (npairs + 1), false, ps));
  ST_TRY(g.counts.grow(sizeof(unsigned long long) * nb, false, ps));
  ST_TRY(g.cells.grow(sizeof(unsigned long long) * nb, false, ps));
  ST_TRY(g.cursor.grow(sizeof(unsigned long long) * nb, false, ps));
  ST_TRY(g.maxlds.grow(sizeof(int32_t) * nb, false, ps));
  ST_TRY(g.tickets.grow(sizeof(unsigned int) * nb, false, ps));
  ST_TRY(g.abort.grow(sizeof(int32_t) * kAbortInts, false, ps));
  ST_TRY(g.tasks.grow(sizeof(uint64_t) * std::max<int64_t>(planned, 1), false, ps));
  ST_TRY(g.descs.grow(sizeof(TaskDesc) * std::max<int64_t>(planned, 1), false, ps));
  ST_TRY(g.td.grow((size_t)g.td_bytes, false, ps));
  ST_TRY(g.dist.grow(sizeof(int32_t) * npairs, false, ps));
  ST_TRY(g.lines_d.grow(sizeof(int64_t) * (npairs + 1), false, ps));
  ST_TRY(g.leaf_d.grow((size_t)npairs, false, ps));
  if (level_align) ST_TRY(g.loff_d.grow(sizeof(int32_t) * npairs, false, ps));
  HIP_TRY(hipMemcpyAsync(g.pairs.p, g.pm.data(), sizeof(PairMeta) * npairs,
                         hipMemcpyHostToDevice, ps));
  HIP_TRY(hipMemcpyAsync(g.blk.p, blk.data(), sizeof(int64_t) * (npairs + 1),
                         hipMemcpyHostToDevice, ps));
  HIP_TRY(hipMemcpyAsync(g.lines_d.p, g.lines.data(), sizeof(int64_t) * (npairs + 1),
                         hipMemcpyHostToDevice, ps));
  HIP_TRY(hipMemcpyAsync(g.tail_lv.p, g.h_tail.data(), nlevels, hipMemcpyHostToDevice, ps));
  HIP_TRY(hipMemcpyAsync(g.leaf_d.p, leaf.data(), (size_t)npairs, hipMemcpyHostToDevice, ps));
  if (level_align)
    HIP_TRY(hipMemcpyAsync(g.loff_d.p, g.loff.data(), sizeof(int32_t) * npairs,
                           hipMemcpyHostToDevice, ps));
  for (hipEvent_t* e : {&g.ev_plan0, &g.ev_plan1, &g.ev_end})
    if (!*e) HIP_TRY(hipEventCreate(e));
  HIP_TRY(hipEventRecord(g.ev_plan0, ps));
  // The leaf lines read only the trees and the arena layout, which are on the
  // device now: a batch's run on its level stream while the plan is built
  // (not when every launch is timed on its own).
  g.leaf_early = false;
  if (lane < 0 && !profile && g.lines[npairs] > 0) {
    hipStream_t ls = lane_stream(lane);
    HIP_TRY(hipStreamWaitEvent(ls, g.ev_plan0, 0));
    LeafArgs la = leaf_args(g);
    la.npairs = npairs;  // g.npairs is set once the plan is complete
    launch_leaf(la, g.lines[npairs], ls);
    HIP_TRY(hipGetLastError());
    ++st.launches;
    g.leaf_early = true;
  }

  PlanArgs pa;
  pa.pool = view();
  pa.pairs = g.pairs.as<PairMeta>();
  pa.blk_start = g.blk.as<int64_t>();
  pa.level_off = level_align ? g.loff_d.as<int32_t>() : nullptr;
  pa.npairs = npairs;
  pa.nlevels = nlevels;
  pa.lchunk_tables = plan_chunk;
  pa.rule = rule;
  // The cost forms with 16 columns per lane hold 169-176 VGPRs (2 waves per
  // SIMD): under a cost model a table gets more waves instead (the 4096-pair
  // batch under a cost model ran 125.5 ms against 131.3 with up to 16).
  if (cost_alphabet_ > 0) pa.rule.max_cpl = std::min(pa.rule.max_cpl, kMaxCostCpl);
  pa.rule_tail = ClassRule{lds_limit, 1, 1 << 30, lds_limit, 16, rule.xl_seq_rows};
  pa.rule_tail.xl_first = rule.xl_first;
  pa.rule_tail.tiny = rule.tiny;
  pa.tail_level = g.tail_lv.as<uint8_t>();
  pa.force_class = force_class;
  pa.bands = size_bands;
  pa.light_per_wave = light_per_wave;
  pa.lane_enable = lane_enable;
  pa.pair_leaf = g.leaf_d.as<uint8_t>();
  pa.counts = g.counts.as<unsigned long long>();
  pa.cells = g.cells.as<unsigned long long>();
  pa.cursor = g.cursor.as<unsigned long long>();
  pa.maxlds = g.maxlds.as<int32_t>();
  pa.tasks = g.tasks.as<uint64_t>();
  pa.descs = g.descs.as<TaskDesc>();
  pa.tickets = g.tickets.as<unsigned int>();
  pa.abort_flag = g.abort.as<int32_t>();
  launch_plan(pa, blk[npairs], ps);
  HIP_TRY(hipGetLastError());
  g.h_counts.assign(nb, 0);
  g.h_maxlds.assign(nb, 0);
  g.h_cells.assign(nb, 0);
  HIP_TRY(hipMemcpyAsync(g.h_counts.data(), g.counts.p, sizeof(unsigned long long) * nb,
                         hipMemcpyDeviceToHost, ps));
  HIP_TRY(hipMemcpyAsync(g.h_maxlds.data(), g.maxlds.p, sizeof(int32_t) * nb,
                         hipMemcpyDeviceToHost, ps));
  HIP_TRY(hipMemcpyAsync(g.h_cells.data(), g.cells.p, sizeof(unsigned long long) * nb,
                         hipMemcpyDeviceToHost, ps));
  HIP_TRY(hipStreamSynchronize(ps));
  HIP_TRY(hipEventRecord(g.ev_plan1, ps));
  // Column-blocked tables: ordinal of each bucket's first task, then (with a
  // second short sync, only in plans that have such tables) the arena
  // layout computed on the device.
  g.xl_first.assign(nb, 0);
  g.nxl = 0;
  for (int L = 0; L < nlevels; ++L) {
    if (g.h_counts[bucket_of(L, kClassXL, 0)]) {
      set_error("run_batch: a table fits no form (nesting height above " +
                std::to_string(kMaxImports - 1) + " or slot storage above " +
                std::to_string(lds_limit) + " bytes even in 256-column blocks)");
      return ERR_TOO_LARGE;
    }
    for (int k = 0; k < kNumXLForms; ++k) {
      const size_t bkt = bucket_of(L, kFirstXLClass + k, 0);
      g.xl_first[bkt] = g.nxl;
      g.nxl += (int64_t)g.h_counts[bkt];
    }
  }
  if (g.nxl > 0) {
    ST_TRY(g.xl_off.grow(sizeof(int64_t) * g.nxl, false, ps));
    ST_TRY(g.xl_total.grow(sizeof(unsigned long long), false, ps));
    XLLayoutArgs xa;
    xa.pool = pa.pool;
    xa.pairs = pa.pairs;
    xa.tasks = pa.tasks;
    xa.counts = pa.counts;
    xa.cursor = pa.cursor;
    xa.nlevels = nlevels;
    xa.off = g.xl_off.as<int64_t>();
    xa.total = g.xl_total.as<unsigned long long>();
    launch_xl_layout(xa, ps);
    HIP_TRY(hipGetLastError());
    unsigned long long words = 0;
    HIP_TRY(hipMemcpyAsync(&words, xa.total, sizeof(words), hipMemcpyDeviceToHost, ps));
    HIP_TRY(hipStreamSynchronize(ps));
    const size_t bytes = sizeof(unsigned long long) * std::max<unsigned long long>(words, 1);
    if (bytes > g.xl_words.cap) {
      ST_TRY(g.xl_words.grow(bytes, false, ps));
      launch_xl_zero(g.xl_words.as<unsigned long long>(), (int64_t)(g.xl_words.cap / 8), ps);
      HIP_TRY(hipGetLastError());
    }
  }
  // Level launches (main stream, and the side streams forked from it) come
  // after everything the plan enqueued, the arena fill included.
  HIP_TRY(hipEventRecord(plan_done_, ps));
  HIP_TRY(hipStreamWaitEvent(lane_stream(lane), plan_done_, 0));
  g.npairs = npairs;
  g.cost_epoch = cost_epoch_;
  g.nlevels = nlevels;
  g.next_level = 0;
  g.level_off = 0;
  st.levels = nlevels;
  return OK;
}

int Context::launch_levels(Group* const* gs, int ng, std::vector<ProfLaunch>* prof, int lane) {
  // Launches of one level are independent, and so are levels of different
  // groups: they fork from the main stream onto up to level_streams streams
  // (largest first) and join before the next tick, so latency-bound launches
  // of few large tables overlap with the rest instead of queueing behind it.
  struct Launch {
    Group* g;
    int cls, bucket;  // bucket: the (level, class) bucket's band 0
    unsigned long long off, cnt, cells;
    int32_t lds;      // largest slot storage over the bands
    int level;
    bool heavy;       // the tables that need more slot storage than light_lds
  };
  std::vector<Launch> lv, leaf;
  for (int gi = 0; gi < ng; ++gi) {
    if (gs[gi]->cost_epoch != cost_epoch_) {
      set_error("launch: the cost model changed after these pairs were planned");
      return ERR_INVALID_ARG;
    }
  }
  for (int gi = 0; gi < ng; ++gi) {
    Group& g = *gs[gi];
    const int L = g.next_level;
    if (L >= g.nlevels) continue;
    // The leaf lines read nothing; the leaf kernel also fills cells of later
    // tables with placeholders, so it runs before the group's first level.
    if (L == 0 && g.lines[g.npairs] > 0 && !g.leaf_early)
      leaf.push_back({&g, kClassLane, -1, 0, (unsigned long long)g.lines[g.npairs],
                      (unsigned long long)g.leaf_cells, 0, -1, false});
    unsigned long long off = g.level_off;
    for (int c = 0; c < kNumClasses; ++c) {
      // The bands of (L, c) are contiguous in the task array, largest first:
      // light tables, then heavy ones (size_band), one launch each.
      for (int half = 0; half < kBands / kRowBands; ++half) {
        unsigned long long cnt = 0, cells = 0;
        int32_t lds = 0;
        for (int bd = half * kRowBands; bd < (half + 1) * kRowBands; ++bd) {
          const size_t b = bucket_of(L, c, bd);
          cnt += g.h_counts[b];
          cells += g.h_cells[b];
          lds = std::max(lds, g.h_maxlds[b]);
        }
        if (cnt) lv.push_back({&g, c, bucket_of(L, c, 0), off, cnt, cells, lds, L, half > 0});
        off += cnt;
      }
    }
    g.level_off = off;
    ++g.next_level;
  }
  // Launches whose time is set by their longest tables first: the heavy
  // ones (few long tables at low occupancy) and those of 4 or more waves
  // per table, each on a stream of its own where there are enough; then the
  // rest by class id, descending. (Heavy launches last ran 110 ms against
  // 101 ms per batch; among the rest, class order measured about 1% faster
  // than ordering by cells.)
  auto latency_bound = [](const Launch& l) {
    return l.heavy || (l.cls >= row_class(2, 0) && l.cls <= kNumRowClasses);
  };
  std::stable_sort(lv.begin(), lv.end(), [&](const Launch& x, const Launch& y) {
    return latency_bound(x) != latency_bound(y) ? latency_bound(x) : x.cls > y.cls;
  });
  if (prof) {
    while (prof_ev_.size() < 2 * (prof->size() + lv.size() + leaf.size())) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      prof_ev_.push_back(e);
    }
  }
  hipStream_t sl[kMaxLevelStreams];
  const int ns = std::max(1, std::min<int>((int)lv.size(), lane_streams(lane, sl)));
  // Leaf launches go first on the main stream; the fork follows them, so
  // every table of the level starts after them.
  const size_t nleaf = leaf.size();
  lv.insert(lv.begin(), leaf.begin(), leaf.end());
  for (size_t i = 0; i < lv.size(); ++i) {
    if (i == nleaf) {
      if (ns > 1) HIP_TRY(hipEventRecord(fork_ev_, sl[0]));
      for (int k = 1; k < ns; ++k) HIP_TRY(hipStreamWaitEvent(sl[k], fork_ev_, 0));
    }
    const Launch& w = lv[i];
    Group& g = *w.g;
    hipStream_t s = i < nleaf ? sl[0] : sl[(i - nleaf) % ns];
    if (prof) HIP_TRY(hipEventRecord(prof_ev_[2 * prof->size()], s));
    if (w.cls == kClassLane) {
      launch_leaf(leaf_args(g), (int64_t)w.cnt, s);
    } else {
      DPArgs da;
      da.pool = view();
      da.pairs = g.pairs.as<PairMeta>();
      da.td = g.td.as<td_t>();
      da.tasks = g.tasks.as<uint64_t>() + w.off;
      da.descs = g.descs.as<TaskDesc>() + w.off;
      da.ntasks = (int64_t)w.cnt;
      da.xl_words = g.xl_words.as<unsigned long long>();
      da.xl_abort = g.abort.as<int32_t>();
      da.xl_off = nullptr;
      da.xl_ticket = nullptr;
      da.xl_epoch = 0;
      da.xl_nbmax = 1;
      da.xl_seq_rows = rule.xl_seq_rows;
      da.ren = cren_.as<int32_t>();
      da.alphabet = cost_alphabet_;
      da.costs = cost_alphabet_ > 0 ? 1 : 0;
      if (w.cls >= kFirstXLClass && w.cls < kFirstXLClass + kNumXLForms) {
        const int bw = xl_bw(w.cls - kFirstXLClass);
        da.xl_off = g.xl_off.as<int64_t>() + g.xl_first[w.bucket];
        da.xl_ticket = g.tickets.as<unsigned int>() + w.bucket;
        // never 0 (a zeroed word) nor 0xFFFFFFFF (the reads' compare value)
        if (++xl_epoch_ == 0xFFFFFFFFu) xl_epoch_ = 1;
        da.xl_epoch = xl_epoch_;
        da.xl_nbmax = (g.maxn + bw - 1) / bw;
      }
      const int lds = std::max(64, (w.lds + 15) & ~15);
      if (!launch_forest(w.cls, da, lds, s)) {
        set_error("run_batch: no kernel for class " + std::to_string(w.cls));
        return ERR_INVALID_ARG;
      }
    }
    if (prof) {
      HIP_TRY(hipEventRecord(prof_ev_[2 * prof->size() + 1], s));
      prof->push_back(ProfLaunch{w.cls, w.level,
                                 w.cls == kClassLane ? g.leaf_tables : (int64_t)w.cnt,
                                 (int64_t)w.cells, &g});
    }
    ++g.st.launches;
  }
  for (int k = 1; k < ns; ++k) {
    HIP_TRY(hipEventRecord(join_ev_[k - 1], sl[k]));
    HIP_TRY(hipStreamWaitEvent(sl[0], join_ev_[k - 1], 0));
  }
  HIP_TRY(hipGetLastError());
  return OK;
}

int Context::reserve_pinned(Group& g, int32_t n) {
  if (g.pin_cap >= n) return OK;
  if (g.h_dist_pin) HIP_TRY(hipHostFree(g.h_dist_pin));
  g.h_dist_pin = nullptr;
  g.pin_cap = 0;
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g.h_dist_pin),
                        sizeof(int32_t) * ((size_t)n + kAbortInts), hipHostMallocDefault));
  g.pin_cap = n;
  return OK;
}

int Context::release_retired() {
  for (void* p : retired_) HIP_TRY(hipFree(p));
  retired_.clear();
  return OK;
}

int Context::finish_async(Group& g, int lane) {
  hipStream_t ls = lane_stream(lane);
  ST_TRY(reserve_pinned(g, g.npairs));
  if (!g.done_ev) HIP_TRY(hipEventCreateWithFlags(&g.done_ev, hipEventDisableTiming));
  launch_gather_dist(g.pairs.as<PairMeta>(), g.td.as<td_t>(), g.npairs, g.dist.as<int32_t>(),
                     ls);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(g.h_dist_pin, g.dist.p, sizeof(int32_t) * g.npairs,
                         hipMemcpyDeviceToHost, ls));
  HIP_TRY(hipMemcpyAsync(g.h_dist_pin + g.npairs, g.abort.p, sizeof(int32_t) * kAbortInts,
                         hipMemcpyDeviceToHost, ls));
  HIP_TRY(hipEventRecord(g.done_ev, ls));
  return OK;
}

int Context::poll(Group& g, bool wait, int32_t* dist) {
  hipError_t e = wait ? hipEventSynchronize(g.done_ev) : hipEventQuery(g.done_ev);
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    return ERR_BUSY;
  }
  HIP_TRY(e);
  std::copy(g.h_dist_pin, g.h_dist_pin + g.npairs, dist);
  if (g.nxl > 0 && g.h_dist_pin[g.npairs]) return xl_timeout(g.h_dist_pin + g.npairs);
  return OK;
}

// The abort record of a column-blocked launch: {flag, claimed, ticket,
// kind << 24 | row, word low, word high, epoch}.
int xl_timeout(const int32_t* rec) {
  set_error("run_batch: a column-blocked table stopped waiting for its neighbour block (ticket " +
            std::to_string(rec[2]) + ", " + ((rec[3] >> 24) ? "import" : "carry") + " of row " +
            std::to_string(rec[3] & 0xFFFFFF) + ", read " + std::to_string((uint32_t)rec[5]) +
            ":" + std::to_string(rec[4]) + ", epoch " + std::to_string((uint32_t)rec[6]) + ")");
  return ERR_TIMEOUT;
}

int Context::sync() {
  HIP_TRY(hipStreamSynchronize(stream_));
  return OK;
}

int Context::drain() {
  HIP_TRY(hipSetDevice(device_));
  HIP_TRY(hipStreamSynchronize(stream_));
  for (hipStream_t sd : side_) HIP_TRY(hipStreamSynchronize(sd));
  HIP_TRY(hipStreamSynchronize(plan_stream_));
  return OK;
}

bool Context::leaf_form(int32_t x, int32_t y) const {
  if (!lane_enable) return false;
  return cost_alphabet_ == 0 ||
         (info_[x].line_stack <= kCostLeafStack && info_[y].line_stack <= kCostLeafStack);
}

LeafArgs Context::leaf_args(const Group& g) const {
  LeafArgs la;
  la.pool = view();
  la.pairs = g.pairs.as<PairMeta>();
  la.tile_start = g.lines_d.as<int64_t>();
  la.npairs = g.npairs;
  la.td = g.td.as<td_t>();
  la.del = cdel_.as<int32_t>();
  la.ins = cins_.as<int32_t>();
  la.ren = cren_.as<int32_t>();
  la.alphabet = cost_alphabet_;
  la.costs = cost_alphabet_ > 0 ? 1 : 0;
  return la;
}

int64_t Context::pair_device_bytes(int32_t x, int32_t y) const {
  const TreeMeta &a = meta_[x], &b = meta_[y];
  const bool lf = leaf_form(x, y);
  const int64_t ka = lf ? a.K - a.K1 : a.K, kb = lf ? b.K - b.K1 : b.K;
  return (2 * (int64_t)a.n * b.n + 3 * kTdPad) * (int64_t)sizeof(td_t) +
         (int64_t)(sizeof(uint64_t) + sizeof(TaskDesc)) * ka * kb;
}

// Everything launch_levels bakes into the launches of a group: the plan's
// counts and LDS sizes, the buffers, the pool arrays and the streams.
std::vector<uint64_t> Context::graph_signature(const Group& g, int lane) const {
  std::vector<uint64_t> s;
  hipStream_t sl[kMaxLevelStreams];
  const int ns = lane_streams(lane, sl);
  s.push_back((uint64_t)ns);
  for (int k = 0; k < ns; ++k) s.push_back((uint64_t)(uintptr_t)sl[k]);
  s.push_back((uint64_t)g.npairs);
  s.push_back((uint64_t)g.nlevels);
  s.push_back((uint64_t)g.lines[g.npairs]);
  for (const void* p : {g.pairs.p, g.td.p, g.tasks.p, g.descs.p, g.lines_d.p})
    s.push_back((uint64_t)(uintptr_t)p);
  const PoolView v = view();
  const uint64_t* pv = reinterpret_cast<const uint64_t*>(&v);
  s.insert(s.end(), pv, pv + sizeof(PoolView) / sizeof(uint64_t));
  s.insert(s.end(), g.h_counts.begin(), g.h_counts.end());
  for (int32_t m : g.h_maxlds) s.push_back((uint64_t)(uint32_t)m);
  s.push_back((uint64_t)cost_alphabet_);
  s.push_back((uint64_t)(uintptr_t)cren_.p);
  s.push_back((uint64_t)g.leaf_early);
  return s;
}

// All levels of a planned group on its lane: as launches, or as a graph.
int Context::run_levels(Group& g, int lane, std::vector<ProfLaunch>* prof) {
  Group* one = &g;
  if (!use_graphs || prof || g.nxl > 0) {
    while (g.next_level < g.nlevels) {
      const int st = launch_levels(&one, 1, prof, lane);
      if (st != OK) return st;
    }
    return OK;
  }
  hipStream_t ls = lane_stream(lane);
  std::vector<uint64_t> sig = graph_signature(g, lane);
  if (!g.graph_exec || sig != g.graph_sig) {
    if (g.graph_exec) HIP_TRY(hipGraphExecDestroy(g.graph_exec));
    if (g.graph) HIP_TRY(hipGraphDestroy(g.graph));
    g.graph_exec = nullptr;
    g.graph = nullptr;
    g.graph_sig.clear();
    HIP_TRY(hipStreamBeginCapture(ls, hipStreamCaptureModeThreadLocal));
    int st = OK;
    while (st == OK && g.next_level < g.nlevels) st = launch_levels(&one, 1, nullptr, lane);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ls, &graph);
    if (st != OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    HIP_TRY(e);
    g.graph = graph;
    HIP_TRY(hipGraphInstantiate(&g.graph_exec, g.graph, nullptr, nullptr, 0));
    g.graph_sig = std::move(sig);
    g.graph_launches = g.st.launches;
  } else {
    g.next_level = g.nlevels;
    g.st.launches = g.graph_launches;
  }
  HIP_TRY(hipGraphLaunch(g.graph_exec, ls));
  return OK;
}

int Context::run_batch(int32_t npairs, const int32_t* xs, const int32_t* ys, int32_t* dist,
                       int32_t* const* td_out, BatchStats* stats) {
  std::lock_guard<std::mutex> lk(mu_);
  const auto t_start = std::chrono::steady_clock::now();
  if (npairs < 0 || (npairs > 0 && (!xs || !ys || !dist))) {
    set_error("run_batch: null argument");
    return ERR_INVALID_ARG;
  }
  if (npairs == 0) {
    if (stats) *stats = BatchStats{};
    return OK;
  }
  const int ng = npairs >= batch_split_min
                     ? std::max(1, std::min<int>(batch_groups, std::min(kMaxLevelStreams, npairs)))
                     : 1;
  // Pair i goes to group i % ng; idx[k] maps a group's pairs back.
  std::vector<std::vector<int32_t>> gx(ng), gy(ng), idx(ng);
  for (int32_t i = 0; i < npairs; ++i) {
    gx[i % ng].push_back(xs[i]);
    gy[i % ng].push_back(ys[i]);
    idx[i % ng].push_back(i);
  }
  std::vector<ProfLaunch> prof;
  int st = OK;
  // Each group: plan (the host waits for its counts while the groups before
  // it run), then every level, on the group's own stream.
  for (int k = 0; k < ng; ++k) {
    Group& g = batch_[k];
    const int lane = ng > 1 ? k : -1;
    plan_share = ng;
    st = plan(g, (int32_t)gx[k].size(), gx[k].data(), gy[k].data(), lane);
    plan_share = 1;
    if (st == OK) st = run_levels(g, lane, profile ? &prof : nullptr);
    if (st != OK) {
      // Launches of this or an earlier group may still be in flight; the next
      // call rewrites these groups' plans, so let them finish first.
      (void)drain();
      return st;
    }
    HIP_TRY(hipEventRecord(g.ev_end, lane_stream(lane)));
  }
  // Distances, abort records and tables of every group, on its stream.
  std::vector<std::vector<int32_t>> gd(ng);
  std::vector<std::array<int32_t, kAbortInts>> aborted(ng);
  for (int k = 0; k < ng; ++k) {
    Group& g = batch_[k];
    hipStream_t s = lane_stream(ng > 1 ? k : -1);
    gd[k].assign(g.npairs, 0);
    aborted[k].fill(0);
    launch_gather_dist(g.pairs.as<PairMeta>(), g.td.as<td_t>(), g.npairs, g.dist.as<int32_t>(),
                       s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(gd[k].data(), g.dist.p, sizeof(int32_t) * g.npairs,
                           hipMemcpyDeviceToHost, s));
    if (g.nxl > 0)
      HIP_TRY(hipMemcpyAsync(aborted[k].data(), g.abort.p, sizeof(int32_t) * kAbortInts,
                             hipMemcpyDeviceToHost, s));
    if (td_out) {
      // Each requested table is widened to int32 in a staging buffer, then
      // copied out (in stream order, so the next widening waits for it).
      int64_t most = 0;
      for (int32_t j = 0; j < g.npairs; ++j)
        if (td_out[idx[k][j]])
          most = std::max(most, (int64_t)meta_[g.xs[j]].n * meta_[g.ys[j]].n);
      if (most > 0) ST_TRY(g.tdw.grow(sizeof(int32_t) * (size_t)most, false, s));
      for (int32_t j = 0; j < g.npairs; ++j) {
        int32_t* out = td_out[idx[k][j]];
        if (!out) continue;
        const int64_t mn = (int64_t)meta_[g.xs[j]].n * meta_[g.ys[j]].n;
        launch_widen_td(g.td.as<td_t>() + g.pm[j].td_off, g.tdw.as<int32_t>(), mn, s);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, g.tdw.p, sizeof(int32_t) * mn, hipMemcpyDeviceToHost, s));
      }
    }
  }
  for (int k = 0; k < ng; ++k) HIP_TRY(hipStreamSynchronize(lane_stream(ng > 1 ? k : -1)));
  for (int k = 0; k < ng; ++k)
    if (aborted[k][0]) return xl_timeout(aborted[k].data());
  for (int k = 0; k < ng; ++k)
    for (size_t j = 0; j < idx[k].size(); ++j) dist[idx[k][j]] = gd[k][j];

  if (profile) {
    for (auto& cp : class_profile) cp = ClassProfile{};
    launch_log.clear();
    for (size_t i = 0; i < prof.size(); ++i) {
      float ms = 0, t0 = 0;
      HIP_TRY(hipEventElapsedTime(&ms, prof_ev_[2 * i], prof_ev_[2 * i + 1]));
      HIP_TRY(hipEventElapsedTime(&t0, prof_ev_[0], prof_ev_[2 * i]));
      ClassProfile& cp = class_profile[prof[i].cls];
      cp.launches += 1;
      cp.ms += ms;
      cp.tables += prof[i].tables;
      cp.cells += prof[i].cells;
      launch_log.push_back(LaunchRecord{prof[i].level, prof[i].cls, prof[i].tables, prof[i].cells,
                                        (double)t0, (double)ms});
    }
  }
  // Plan time: the groups' plans added up; DP time: from the end of the
  // first plan to the end of the last group's levels (later plans overlap).
  BatchStats out;
  double dp_ms = 0;
  for (int k = 0; k < ng; ++k) {
    const Group& g = batch_[k];
    float pms = 0, dms = 0;
    HIP_TRY(hipEventElapsedTime(&pms, g.ev_plan0, g.ev_plan1));
    HIP_TRY(hipEventElapsedTime(&dms, batch_[0].ev_plan1, g.ev_end));
    out.plan_ms += pms;
    dp_ms = std::max(dp_ms, (double)dms);
    out.cells += g.st.cells;
    out.tables += g.st.tables;
    out.levels = std::max(out.levels, g.st.levels);
    out.launches += g.st.launches;
    out.xl_stale += aborted[k][8];
    out.xl_waits += aborted[k][9];
  }
  out.dp_ms = dp_ms;
  out.total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (stats) *stats = out;
  return OK;
}

// Bytes plan() carves for these pairs, every buffer rounded to 256 bytes as
// the arena does (the column-blocked words, sized on the device, come on top).
int64_t Context::noalloc_bytes(int32_t npairs, const int32_t* xs, const int32_t* ys) const {
  auto r = [](int64_t b) { return (b + 255) & ~int64_t{255}; };
  int64_t td = 0, planned = 0;
  int32_t nlevels = 0;
  for (int32_t i = 0; i < npairs; ++i) {
    const TreeMeta &a = meta_[xs[i]], &b = meta_[ys[i]];
    const bool le = leaf_form(xs[i], ys[i]);
    td += 2 * (int64_t)a.n * b.n + 3 * kTdPad;
    planned += (int64_t)(le ? a.K - a.K1 : a.K) * (le ? b.K - b.K1 : b.K);
    nlevels = std::max(nlevels, a.kdmax + b.kdmax + 1);
  }
  const int64_t nb = (int64_t)nlevels * kNumClasses * kBands;
  planned = std::max<int64_t>(planned, 1);
  return r(sizeof(PairMeta) * npairs) + r(nlevels) + 2 * r(8 * (int64_t)(npairs + 1)) +
         3 * r(8 * nb) + 2 * r(4 * nb) + r(32) + r(8 * planned) + r(64 * planned) +
         r((int64_t)sizeof(td_t) * td) +
         r(4 * (int64_t)npairs) + (level_align ? r(4 * (int64_t)npairs) : 0) + r(npairs);
}

int Context::run_batch_noalloc(int32_t npairs, const int32_t* xs, const int32_t* ys,
                               int32_t* dist) {
  std::lock_guard<std::mutex> lk(mu_);
  if (npairs < 0 || (npairs > 0 && (!xs || !ys || !dist))) {
    set_error("batch_noalloc: null argument");
    return ERR_INVALID_ARG;
  }
  if (npairs == 0) return OK;
  if (!na_arena_.base) {
    set_error("batch_noalloc: the context has no device scratch array");
    return ERR_DEVICE;
  }
  for (int32_t i = 0; i < npairs; ++i) {
    if (xs[i] < 0 || ys[i] < 0 || xs[i] >= (int32_t)meta_.size() ||
        ys[i] >= (int32_t)meta_.size()) {
      set_error("batch_noalloc: tree id out of range in pair " + std::to_string(i));
      return ERR_NOT_FOUND;
    }
  }
  const int64_t need = noalloc_bytes(npairs, xs, ys);
  if (need > (int64_t)na_arena_.cap) {
    set_error("batch_noalloc: the pairs need " + std::to_string(need) +
              " bytes of device scratch; the array holds " + std::to_string(na_arena_.cap));
    return ERR_TOO_LARGE;
  }
  HIP_TRY(hipSetDevice(device_));
  // Every buffer is carved afresh from the start of the array.
  na_arena_.used = 0;
  na_.for_each_buf([](DevBuf& b) {
    b.p = nullptr;
    b.cap = 0;
  });
  // No graph (instantiating one allocates) and no per-launch events.
  const int32_t graphs = use_graphs, prof = profile;
  use_graphs = 0;
  profile = 0;
  int st = plan(na_, npairs, xs, ys, -1);
  if (st == OK) st = run_levels(na_, -1, nullptr);
  use_graphs = graphs;
  profile = prof;
  if (st != OK) {
    (void)drain();  // launches made before the failure finish first
    return st;
  }
  hipStream_t s = lane_stream(-1);
  int32_t rec[kAbortInts] = {};
  launch_gather_dist(na_.pairs.as<PairMeta>(), na_.td.as<td_t>(), npairs,
                     na_.dist.as<int32_t>(), s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(dist, na_.dist.p, sizeof(int32_t) * npairs, hipMemcpyDeviceToHost, s));
  if (na_.nxl > 0)
    HIP_TRY(hipMemcpyAsync(rec, na_.abort.p, sizeof(rec), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (rec[0]) return xl_timeout(rec);
  return OK;
}

}  // namespace xted
