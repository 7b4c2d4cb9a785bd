// Continuous-batching server (see server.hpp).
#include "server.hpp"

namespace xted {

Server::Server(Context* ctx, int64_t hbm_budget, int32_t max_batch)
    : ctx_(ctx), budget_(hbm_budget), max_batch_(max_batch) {
  worker_ = std::thread([this]() { loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  if (worker_.joinable()) worker_.join();
}

int Server::submit(int32_t x, int32_t y, uint64_t tag) {
  if (x < 0 || y < 0 || x >= ctx_->pool_size() || y >= ctx_->pool_size()) {
    set_error("server submit: tree id out of range");
    return ERR_NOT_FOUND;
  }
  const int64_t bytes = 8 * (int64_t)ctx_->tree_meta(x).n * ctx_->tree_meta(y).n;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) return ERR_CANCELLED;
    pending_.push_back(Req{x, y, tag, Clock::now(), bytes});
  }
  cv_work_.notify_one();
  return OK;
}

int Server::poll(int32_t cap, int32_t timeout_ms, std::vector<ServerResult>& out) {
  std::unique_lock<std::mutex> lk(mu_);
  if (done_.empty() && timeout_ms > 0)
    cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&]() { return !done_.empty(); });
  while (!done_.empty() && (int32_t)out.size() < cap) {
    out.push_back(done_.front());
    done_.pop_front();
  }
  return OK;
}

void Server::loop() {
  for (;;) {
    std::vector<Req> batch, rejected;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&]() { return stop_ || !pending_.empty(); });
      if (stop_ && pending_.empty()) return;
      int64_t used = 0;
      while (!pending_.empty() && (int32_t)batch.size() < max_batch_) {
        const Req& r = pending_.front();
        if (r.bytes > budget_) {
          rejected.push_back(r);
          pending_.pop_front();
          continue;
        }
        if (used + r.bytes > budget_) break;
        used += r.bytes;
        batch.push_back(r);
        pending_.pop_front();
      }
    }
    const auto t_admit = Clock::now();
    std::vector<int32_t> xs(batch.size()), ys(batch.size()), dist(batch.size(), -1);
    for (size_t i = 0; i < batch.size(); ++i) {
      xs[i] = batch[i].x;
      ys[i] = batch[i].y;
    }
    int st = OK;
    if (!batch.empty())
      st = ctx_->run_batch((int32_t)batch.size(), xs.data(), ys.data(), dist.data(), nullptr,
                           nullptr);
    const auto t_done = Clock::now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (size_t i = 0; i < batch.size(); ++i) {
        ServerResult r;
        r.tag = batch[i].tag;
        r.status = st;
        r.dist = st == OK ? dist[i] : -1;
        r.latency_ms = std::chrono::duration<double, std::milli>(t_done - batch[i].t_submit).count();
        r.queue_ms = std::chrono::duration<double, std::milli>(t_admit - batch[i].t_submit).count();
        done_.push_back(r);
      }
      for (const Req& q : rejected) {
        ServerResult r{q.tag, ERR_TOO_LARGE, -1,
                       std::chrono::duration<double, std::milli>(t_done - q.t_submit).count(), 0.0};
        done_.push_back(r);
      }
    }
    cv_done_.notify_all();
  }
}

}  // namespace xted
