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

// Thread-local message for the last failing call on this thread.
void set_error(const std::string& msg);
const char* last_error();

// Streams (the context's own plus side streams) that the launches of one
// level are spread over; four matches the hardware queues HIP gives a
// process by default.
constexpr int kMaxLevelStreams = 4;

struct BatchStats {
  double plan_ms = 0;      // device plan build incl. the count read-back
  double dp_ms = 0;        // level loop (all forest kernels)
  double total_ms = 0;     // whole call, host inputs -> host distances
  int64_t cells = 0;       // interior forest cells (sum of S_x * S_y)
  int64_t tables = 0;      // keyroot-pair tables
  int32_t levels = 0;
  int32_t launches = 0;    // forest kernel launches
};

// Growable device buffer. grow() keeps contents when `keep` is set.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
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

struct HostTreeInfo {
  int32_t n = 0;
  int64_t ksum = 0;  // S = sum of keyroot subtree sizes
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
  int32_t pool_size() const { return (int32_t)meta_.size(); }
  const TreeMeta& tree_meta(int32_t id) const { return meta_[id]; }
  const HostTreeInfo& tree_info(int32_t id) const { return info_[id]; }
  // Copies the device-computed orl, keyroots (descending) and keyroot depths.
  int tree_arrays(int32_t id, int32_t* orl, int32_t* kr, int32_t* K, int32_t* kd);

  // Evaluation -------------------------------------------------------------
  // Distances of pairs (xs[i], ys[i]); when td_out is given, td_out[i] (may be
  // null per pair) receives the m x n tree-distance table.
  int run_batch(int32_t npairs, const int32_t* xs, const int32_t* ys, int32_t* dist,
                int32_t* const* td_out, BatchStats* stats);

  // Tuning / test hooks -----------------------------------------------------
  ClassRule rule{kLdsPerWave, 1, 1 << 30, kLdsMaxTable, 8};
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
  ClassProfile class_profile[kNumClasses];

  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  hipStream_t side_[kMaxLevelStreams - 1] = {};
  hipEvent_t ev_[4] = {};
  hipEvent_t fork_ev_ = nullptr, join_ev_[kMaxLevelStreams - 1] = {};
  std::mutex mu_;
  // pool
  std::vector<TreeMeta> meta_;
  std::vector<HostTreeInfo> info_;
  int64_t pool_nodes_ = 0;  // used entries (sum of n+1)
  DevBuf parent_, label_, orl_, kdep_, kdn_, lrank_, kr_, krkd_, krsize_, krleaf_, trees_;
  DevBuf skr_, skd_, ssize_, sleaf_, spre_, kdoff_, rtop_;
  DevBuf prep_scratch_, prep_off_;
  // batch
  DevBuf pairs_, blk_, counts_, cells_, cursor_, maxlds_, tasks_, td_, dist_;
  DevBuf lines_;
  // column-blocked tables: launch tickets, abort flag, per-task arena offsets,
  // the arena of cross-workgroup words and the tag of the latest launch
  DevBuf tickets_, abort_, xl_off_, xl_total_, xl_words_, tail_lv_;
  std::vector<uint8_t> h_tail_;
  uint32_t xl_epoch_ = 0;
  std::vector<unsigned long long> h_counts_, h_cells_;
  std::vector<int32_t> h_maxlds_;  // largest slot storage per (level, class) bucket
  std::vector<hipEvent_t> prof_ev_;
  int grow_pool(int64_t nodes);
  PoolView view() const;
};

}  // namespace xted
