// Continuous-batching TED server over a context's device-resident pool.
//
// Queries (x_id, y_id) are queued on submit. A server thread repeatedly
// admits queued queries in arrival order while their tables fit the HBM
// budget (TD + TD^T = 8*m*n bytes each) and the batch cap, evaluates the
// admitted set as one device batch (their tables interleave in the same
// level launches), and publishes the results. Queries that arrive while a
// batch is on the device join the next batch as soon as it forms.
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
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
  struct Req {
    int32_t x, y;
    uint64_t tag;
    Clock::time_point t_submit;
    int64_t bytes;
  };
  void loop();
  Context* ctx_;
  int64_t budget_;
  int32_t max_batch_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Req> pending_;
  std::deque<ServerResult> done_;
  bool stop_ = false;
  std::thread worker_;
};

}  // namespace xted
