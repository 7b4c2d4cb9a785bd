// Host engine over the HIP kernels: a device-resident tree pool, batched
// pair evaluation (plan on device, level-by-level launches), status codes.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "device.hpp"
#include "trees.hpp"

namespace xted {

struct LeafArgs;  // kernels.hpp

// Thread-local message for the last failing call on this thread.
void set_error(const std::string& msg);
const char* last_error();

// Streams (the context's own plus side streams) that the launches of one
// level are spread over; four matches the hardware queues HIP gives a
// process by default.
constexpr int kMaxLevelStreams = 4;

// Device bytes of the fixed array the no-allocation batch path carves its
// buffers from (one per device, held until the process ends).
constexpr size_t kNoAllocBytes = size_t{64} << 20;

struct BatchStats {
  double plan_ms = 0;      // device plan build incl. the count read-back
  double dp_ms = 0;        // level loop (all forest kernels)
  double total_ms = 0;     // whole call, host inputs -> host distances
  int64_t cells = 0;       // interior forest cells (sum of S_x * S_y)
  int64_t tables = 0;      // keyroot-pair tables
  int32_t levels = 0;
  int32_t launches = 0;    // forest kernel launches
  // Column-blocked launches: prefetched cross-workgroup reads that were
  // stale and had to be polled, and pipeline checks that had to wait.
  int64_t xl_stale = 0, xl_waits = 0;
};

// A fixed device region that buffers are carved from in order (no
// allocation): the no-allocation batch path takes every buffer from it.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
};

// Growable device buffer. grow() keeps contents when `keep` is set. A buffer
// bound to an arena never allocates: grow() carves the next piece of the
// arena (contents are not kept) or fails with ERR_TOO_LARGE, and release()
// only forgets the piece. A buffer with a retire list hands its old memory
// to that list when it grows instead of freeing it (hipFree waits for the
// whole device): the owner frees the list when nothing runs.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  bool fine = false;  // fine-grained allocation (hipDeviceMallocFinegrained)
  Arena* arena = nullptr;
  std::vector<void*>* retire = nullptr;
  int grow(size_t bytes, bool keep, hipStream_t s);
  void release();
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// Per granularity class totals of the last batch run with profiling on.
struct ClassProfile {
  int64_t launches = 0;
  int64_t tables = 0;
  int64_t cells = 0;
  double ms = 0;  // summed device time of the class's launches
};

// One planned set of pairs: its plan (tasks bucketed by level and class),
// its TD arena and its position in the level loop. A batch is one group; the
// server keeps several in flight, each at its own level.
struct Group {
  int32_t npairs = 0, nlevels = 0, maxn = 1;
  uint64_t cost_epoch = 0;          // cost model the group was planned under
  bool leaf_early = false;          // the leaf lines were launched while planning
  int32_t next_level = 0;           // next level to launch
  unsigned long long level_off = 0; // tasks of the levels already launched
  int64_t nxl = 0, leaf_tables = 0, leaf_cells = 0, td_bytes = 0;
  std::vector<int32_t> xs, ys;
  std::vector<PairMeta> pm;
  std::vector<int64_t> lines, xl_first;  // lines: prefix of leaf-kernel tiles per pair
  std::vector<int32_t> loff;             // per pair: level offset (Context::level_align)
  std::vector<unsigned long long> h_counts, h_cells;
  std::vector<int32_t> h_maxlds;
  std::vector<uint8_t> h_tail;
  DevBuf pairs, blk, counts, cells, cursor, maxlds, tickets, tasks, descs, td, dist, lines_d,
      tail_lv, xl_off, xl_total, loff_d, leaf_d;
  DevBuf tdw;  // requested tables widened to int32 on their way out
  DevBuf abort, xl_words;
  // finish_async: distances then the abort flag land in pinned memory
  int32_t* h_dist_pin = nullptr;
  int32_t pin_cap = 0;
  hipEvent_t done_ev = nullptr;
  hipEvent_t ev_plan0 = nullptr, ev_plan1 = nullptr, ev_end = nullptr;  // batch timing
  // The group's level launches captured as a graph, and the launch inputs
  // (plan counts, buffers, streams) it was captured for.
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint64_t> graph_sig;
  int32_t graph_launches = 0;
  BatchStats st;
  Group() = default;
  Group(const Group&) = delete;
  Group& operator=(const Group&) = delete;
  ~Group() { release(); }
  void release();
  // Every device buffer of the group.
  template <class F>
  void for_each_buf(F f) {
    for (DevBuf* b : {&pairs, &blk, &counts, &cells, &cursor, &maxlds, &tickets, &abort, &tasks,
                      &descs, &td, &dist, &lines_d, &tail_lv, &xl_off, &xl_total, &xl_words,
                      &loff_d, &leaf_d, &tdw})
      f(*b);
  }
};

// One timed launch of the last profiled call: level (-1 for the leaf
// kernel), class, tables, interior cells, start (ms after the call's first
// timed launch started) and device duration.
struct LaunchRecord {
  int32_t level, cls;
  int64_t tables, cells;
  double start_ms, ms;
};

// One timed forest launch: its class, level (-1 for the leaf kernel), tables,
// interior cells and the group it belongs to.
struct ProfLaunch {
  int cls, level;
  int64_t tables, cells;
  const Group* g;
};

struct HostTreeInfo {
  int32_t n = 0;
  int64_t ksum = 0;    // S = sum of keyroot subtree sizes
  int32_t minlab = 0, maxlab = 0;  // smallest and largest label id
  int64_t sdel = 0, sins = 0;  // cost mode: deleting / inserting the whole tree
  int32_t line_stack = 0;      // line_stack_depth of the tree
};

class Context {
 public:
  explicit Context(int device);
  ~Context();
  int init();

  // Tree pool -------------------------------------------------------------
  // Uploads ntrees preorder trees (sizes[t] nodes each, arrays concatenated)
  // and runs the device preprocessing. Writes the new tree ids.
  int pool_add(int32_t ntrees, const int32_t* sizes, const int32_t* parents,
               const int32_t* labels, int32_t* ids);
  int pool_clear();
  // Non-unit costs per label (labels in [0, alphabet), each cost in
  // [0, 32767]): del, ins (alphabet entries) and ren (alphabet x alphabet,
  // row = x label). del == nullptr restores unit costs.
  int set_costs(int32_t alphabet, const int32_t* del, const int32_t* ins, const int32_t* ren);
  bool cost_mode() const { return cost_alphabet_ > 0; }
  uint64_t cost_epoch() const { return cost_epoch_; }  // changes with every set_costs
  // Cost of deleting (del) or inserting every node of a tree under the cost
  // model: the distance between it and the empty tree. -1 for a label outside
  // the alphabet.
  int64_t whole_tree_cost(const int32_t* labels, int32_t n, bool del) const;
  int32_t pool_size() const { return (int32_t)meta_.size(); }
  // Device bytes a batch holds for pair (x, y) of the pool: its TD and TD^T
  // (8*m*n plus padding) and its plan arrays, a task word and a 64-byte
  // descriptor per planned keyroot-pair table (tables of two non-leaf
  // keyroots with unit costs, every table under a cost model). Both ids
  // must be valid; the caller holds mutex().
  int64_t pair_device_bytes(int32_t x, int32_t y) const;
  // Waits for everything launched on every stream of the context (the level
  // streams and the plan stream), so that a group's buffers can be reused
  // after a failed launch without in-flight kernels reading rewritten plans.
  int drain();
  const TreeMeta& tree_meta(int32_t id) const { return meta_[id]; }
  const HostTreeInfo& tree_info(int32_t id) const { return info_[id]; }
  // Copies the device-computed orl, keyroots (descending) and keyroot depths.
  int tree_arrays(int32_t id, int32_t* orl, int32_t* kr, int32_t* K, int32_t* kd);

  // Evaluation -------------------------------------------------------------
  // Distances of pairs (xs[i], ys[i]); when td_out is given, td_out[i] (may be
  // null per pair) receives the m x n tree-distance table.
  // run_batch without any device allocation: every buffer of the batch (plan,
  // TD arena, distances) is carved from the fixed device array of
  // kNoAllocBytes this library holds per device (its address is looked up
  // once, at init); launches go one by one (no graph is instantiated), the
  // distances are copied to the caller's memory directly. Pairs that do not
  // fit are refused with ERR_TOO_LARGE. Calls that overlap on one device
  // (from different contexts) share the array and race.
  int run_batch_noalloc(int32_t npairs, const int32_t* xs, const int32_t* ys, int32_t* dist);
  int run_batch(int32_t npairs, const int32_t* xs, const int32_t* ys, int32_t* dist,
                int32_t* const* td_out, BatchStats* stats);

  // The steps of run_batch, for callers that interleave several groups (the
  // server). The caller holds mutex() across them.
  // Plans pairs (xs[i], ys[i]) into g: device plan, one sync for the counts.
  // lane: -1 = all streams (a batch); 0 .. kMaxLevelStreams-1 = server lanes,
  // one stream each, whose ticks run independently of each other.
  int plan(Group& g, int32_t npairs, const int32_t* xs, const int32_t* ys, int lane = -1);
  // One tick: launches the next level of every group in gs (groups past
  // their last level are skipped) and advances them. Asynchronous.
  int launch_levels(Group* const* gs, int ng, std::vector<ProfLaunch>* prof, int lane = -1);
  std::mutex& mutex() { return mu_; }
  int sync();  // waits for everything launched on the context's streams
  // Enqueues the distance gather of a group whose levels are all launched;
  // poll() returns ERR_BUSY until it has landed (wait = block for it).
  int finish_async(Group& g, int lane = -1);
  // Pinned host memory for the distances (and the column-blocked record) of
  // up to n pairs of g.
  int reserve_pinned(Group& g, int32_t n);
  // Device memory that grown group buffers handed over (DevBuf::retire),
  // and its release; the caller holds mutex() and nothing of theirs runs.
  std::vector<void*>* retire_list() { return &retired_; }
  int release_retired();
  hipStream_t lane_stream(int lane) const;  // where a lane's ticks end
  int lane_streams(int lane, hipStream_t* out) const;
  int poll(Group& g, bool wait, int32_t* dist);

  // Tuning / test hooks -----------------------------------------------------
  ClassRule rule{kLdsPerWave, 1, 1 << 30, kLdsMaxTable, 8, kXLSeqRows};
  int32_t lds_limit = 0;  // device's dynamic LDS per workgroup for the slots
  int32_t force_class = -1;
  int32_t lane_enable = 1;
  int32_t plan_chunk = 2048;  // tables per plan block
  int32_t profile = 0;        // time every forest launch with events
  int32_t level_streams = kMaxLevelStreams;  // streams one level's launches spread over
  // Levels reached by at most this many pairs of a batch take the latency
  // rule (fewest waves, up to 16 columns per lane, LDS up to the device
  // limit); 0 turns it off.
  int32_t tail_pairs = 32;
  // A batch of at least batch_split_min pairs is cut into batch_groups groups
  // (pair i in group i mod batch_groups), each with its own plan and its own
  // stream: their levels run independently, so one group's level barriers
  // and last tables overlap the other groups' work, and a group's plan
  // overlaps the levels of the groups planned before it. Off by default: on
  // the 4096-pair batch, 2 and 4 groups (one stream each) ran 16% and 65%
  // slower than one group whose levels fork over 4 streams.
  int32_t batch_groups = 1;
  int32_t batch_split_min = 256;
  // Groups a batch is cut into while planning one of them: a level counts
  // as a tail level by the pairs of the whole batch that reach it
  // (estimated as the group's count times this).
  int32_t plan_share = 1;
  // Level launches of a batch without column-blocked tables run as a
  // captured graph (re-launched while the plan's launches stay the same);
  // not while profiling, which times each launch with events.
  int32_t use_graphs = 1;
  // Where a pair's levels sit among the batch's: 0 = every pair starts at
  // level 0, 1 = every pair ends at the batch's last level, 2 = midway. A
  // pair's own levels keep their order, so each is a valid schedule; the
  // choice moves shallow pairs' tables into levels whose time is set by the
  // few large tables of the deepest pairs.
  int32_t level_align = 0;
  // 1: the tables of a launch start largest first (size bands); 2: and the
  // tables that need more slot storage than light_lds take launches of
  // their own; 0: plan order.
  int32_t size_bands = 1;
  // Mode 2: slot bytes per wave above which a row-scan table is heavy.
  int32_t light_per_wave = kLightPerWave;
  ClassProfile class_profile[kNumClasses];
  std::vector<LaunchRecord> launch_log;  // every timed launch of the last profiled call

  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  hipStream_t plan_stream_ = nullptr;
  hipStream_t side_[kMaxLevelStreams - 1] = {};
  hipEvent_t ev_[4] = {};
  hipEvent_t fork_ev_ = nullptr, join_ev_[kMaxLevelStreams - 1] = {};
  hipEvent_t plan_done_ = nullptr;
  std::mutex mu_;
  // pool
  std::vector<TreeMeta> meta_;
  std::vector<HostTreeInfo> info_;
  int64_t pool_nodes_ = 0;  // used entries (sum of n+1)
  DevBuf parent_, label_, orl_, kdep_, kdn_, lrank_, kr_, krkd_, krsize_, krleaf_, trees_;
  DevBuf skr_, skd_, ssize_, sleaf_, spre_, kdoff_, rtop_, crec_, rrec_, ccrec_;
  DevBuf prep_scratch_, prep_off_;
  Group batch_[kMaxLevelStreams];  // the groups run_batch reuses
  Group na_;                        // run_batch_noalloc's group, bound to na_arena_
  std::vector<void*> retired_;      // see retire_list()
  Arena na_arena_;                  // this device's fixed array (kNoAllocBytes)
  int64_t noalloc_bytes(int32_t npairs, const int32_t* xs, const int32_t* ys) const;
  // Whether the leaf lines of pair (x, y) are written in closed form (the
  // leaf kernel) rather than planned as tables: lane_enable, and under a
  // cost model both trees within the cost form's stack.
  bool leaf_form(int32_t x, int32_t y) const;
  LeafArgs leaf_args(const Group& g) const;
  uint32_t xl_epoch_ = 0;  // tag of the latest column-blocked launch
  std::vector<hipEvent_t> prof_ev_;
  int grow_pool(int64_t nodes);
  int cost_prefix(int32_t first, int32_t ntrees);  // pdel_ / pins_ and totals of those trees
  int32_t cost_alphabet_ = 0;                       // 0: unit costs
  std::vector<int32_t> hdel_, hins_;                // host copies of the per-label costs
  uint64_t cost_epoch_ = 0;                         // bumped by every set_costs
  DevBuf cdel_, cins_, cren_, pdel_, pins_, ctot_;
  std::vector<uint64_t> graph_signature(const Group& g, int lane) const;
  int run_levels(Group& g, int lane, std::vector<ProfLaunch>* prof);
  PoolView view() const;
};

}  // namespace xted
