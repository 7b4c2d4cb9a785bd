// Continuous-batching TED server over a context's device-resident pool.
//
// Queries (x_id, y_id) are queued on submit. The server thread works in
// ticks. Each tick it admits queued queries in arrival order while their
// device bytes fit what is left of the HBM budget (TD + TD^T, 8*m*n bytes,
// plus the plan's 72 bytes per planned keyroot-pair table)
// and plans them as a new group; then it launches the next level of every
// group in flight together (levels of different groups are independent, so
// their tables share the same launches' streams); a group whose last level
// has run publishes its results and frees its share of the budget. So a
// query that arrives while others are on the device starts within a tick.
// A tick lasts as long as its slowest table, so groups run in lanes with
// ticks of their own, one stream each: light pairs (the smaller tree at
// most kLightSide nodes and at most kLightLevels levels) in lane 0, heavy
// ones in the least-loaded of lanes 1..3; a light pair never waits for a
// tick of a heavy one, and heavy groups only share ticks within a lane.
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace xted {

struct ServerResult {
  uint64_t tag;
  int32_t status;
  int32_t dist;
  double latency_ms;
  double queue_ms;
};

class Server {
 public:
  Server(Context* ctx, int64_t hbm_budget, int32_t max_batch);
  ~Server();
  int submit(int32_t x, int32_t y, uint64_t tag);
  int poll(int32_t cap, int32_t timeout_ms, std::vector<ServerResult>& out);

 private:
  using Clock = std::chrono::steady_clock;
  static constexpr int32_t kLightSide = 1024;
  static constexpr int32_t kLightLevels = 96;
  struct Req {
    int32_t x, y;
    uint64_t tag;
    Clock::time_point t_submit;
    int64_t bytes;
  };
  static constexpr int kLanes = kMaxLevelStreams;
  struct Active {
    std::unique_ptr<Group> g;
    std::vector<Req> reqs;
    int64_t bytes;
    Clock::time_point t_admit;
  };
  void loop();
  void admit(int lane);
  bool tick(int lane);  // false when nothing was launched
  void publish(const std::vector<Req>& reqs, int status, const int32_t* dist,
               Clock::time_point t_admit);
  void recycle(std::unique_ptr<Group> g);  // error paths: drain, then keep the buffers
  Context* ctx_;
  int64_t budget_;
  int32_t max_batch_;
  int32_t max_groups_ = 16;  // per lane
  // Loop-thread state per lane: groups in flight and the events of the
  // lane's enqueued ticks.
  struct Lane {
    std::vector<Active> active;
    std::deque<hipEvent_t> ticks;
  };
  Lane lanes_[kLanes];
  std::vector<Active> completing_;             // gathers enqueued, not landed
  std::vector<hipEvent_t> free_ev_;
  std::vector<std::unique_ptr<Group>> spare_;  // finished groups, buffers kept
  int64_t inflight_bytes_ = 0;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Req> pending_[2];  // light, heavy
  std::deque<ServerResult> done_;
  bool stop_ = false;
  std::thread worker_;
};

}  // namespace xted
