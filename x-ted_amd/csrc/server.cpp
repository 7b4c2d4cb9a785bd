// Continuous-batching server (see server.hpp).
#include "server.hpp"

#include <cstdio>

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
  std::lock_guard<std::mutex> lk(ctx_->mutex());
  (void)ctx_->drain();
  (void)ctx_->release_retired();
}

int Server::submit(int32_t x, int32_t y, uint64_t tag) {
  // The pool may grow concurrently (xted_pool_add): read it under the
  // context's lock and keep copies.
  int64_t bytes;
  int lane;
  {
    std::lock_guard<std::mutex> lk(ctx_->mutex());
    if (x < 0 || y < 0 || x >= ctx_->pool_size() || y >= ctx_->pool_size()) {
      set_error("server submit: tree id out of range");
      return ERR_NOT_FOUND;
    }
    const TreeMeta a = ctx_->tree_meta(x), b = ctx_->tree_meta(y);
    bytes = ctx_->pair_device_bytes(x, y);
    lane = (std::min(a.n, b.n) > kLightSide || a.kdmax + b.kdmax + 1 > kLightLevels);
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) return ERR_CANCELLED;
    pending_[lane].push_back(Req{x, y, tag, Clock::now(), bytes});
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

void Server::publish(const std::vector<Req>& reqs, int status, const int32_t* dist,
                     Clock::time_point t_admit) {
  const auto t_done = Clock::now();
  if (status != OK && status != ERR_TOO_LARGE && !reqs.empty())
    std::fprintf(stderr, "xted server: %zu queries failed (status %d): %s\n", reqs.size(), status,
                 last_error());
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t i = 0; i < reqs.size(); ++i) {
      ServerResult r;
      r.tag = reqs[i].tag;
      r.status = status;
      r.dist = (status == OK && dist) ? dist[i] : -1;
      r.latency_ms = std::chrono::duration<double, std::milli>(t_done - reqs[i].t_submit).count();
      r.queue_ms = std::max(
          0.0, std::chrono::duration<double, std::milli>(t_admit - reqs[i].t_submit).count());
      done_.push_back(r);
    }
  }
  cv_done_.notify_all();
}

// Admission for one lane: one new group per tick, from the lane's queue
// head, within what is left of the budget.
void Server::admit(int lane) {
  Lane& L = lanes_[lane];
  std::vector<Req> batch, rejected;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if ((int32_t)L.active.size() >= max_groups_) return;
    if (lane >= 1) {
      // heavy queries go to the heavy lane with the fewest groups
      for (int k = 1; k < kLanes; ++k)
        if (lanes_[k].active.size() < L.active.size()) return;
    }
    std::deque<Req>& q = pending_[lane == 0 ? 0 : 1];
    int64_t used = inflight_bytes_;
    while (!q.empty() && (int32_t)batch.size() < max_batch_) {
      const Req& r = q.front();
      if (r.bytes > budget_) {
        rejected.push_back(r);
        q.pop_front();
        continue;
      }
      if (used + r.bytes > budget_) break;
      used += r.bytes;
      batch.push_back(r);
      q.pop_front();
    }
  }
  if (!rejected.empty()) publish(rejected, ERR_TOO_LARGE, nullptr, Clock::now());
  if (batch.empty()) return;
  std::unique_ptr<Group> g;
  if (!spare_.empty()) {
    g = std::move(spare_.back());
    spare_.pop_back();
  } else {
    // A new group: buffers that outgrow themselves go to the context's
    // retire list (hipFree would wait for every lane's work), and the
    // pinned result buffer holds the largest group from the start.
    g.reset(new Group());
    g->for_each_buf([this](DevBuf& b) { b.retire = ctx_->retire_list(); });
    std::lock_guard<std::mutex> lk(ctx_->mutex());
    if (ctx_->reserve_pinned(*g, max_batch_) != OK) {
      publish(batch, ERR_OUT_OF_MEMORY, nullptr, Clock::now());
      return;
    }
  }
  auto fill = [&](std::vector<int32_t>& xs, std::vector<int32_t>& ys) {
    xs.resize(batch.size());
    ys.resize(batch.size());
    int64_t bytes = 0;
    for (size_t i = 0; i < batch.size(); ++i) {
      xs[i] = batch[i].x;
      ys[i] = batch[i].y;
      bytes += batch[i].bytes;
    }
    return bytes;
  };
  std::vector<int32_t> xs, ys;
  int64_t bytes = fill(xs, ys);
  const auto t_admit = Clock::now();
  int st;
  {
    std::lock_guard<std::mutex> lk(ctx_->mutex());
    st = ctx_->plan(*g, (int32_t)batch.size(), xs.data(), ys.data(), lane);
    if (st == ERR_TOO_LARGE && batch.size() > 1) {
      // One pair has a table no form holds: plan pair by pair, reject those,
      // and re-plan the rest together.
      std::vector<Req> keep, drop;
      for (const Req& r : batch)
        (ctx_->plan(*g, 1, &r.x, &r.y, lane) == OK ? keep : drop).push_back(r);
      publish(drop, ERR_TOO_LARGE, nullptr, Clock::now());
      batch = std::move(keep);
      bytes = fill(xs, ys);
      st = batch.empty() ? ERR_TOO_LARGE
                         : ctx_->plan(*g, (int32_t)batch.size(), xs.data(), ys.data(), lane);
    }
  }
  if (st != OK) {
    publish(batch, st, nullptr, t_admit);
    recycle(std::move(g));
    return;
  }
  inflight_bytes_ += bytes;
  L.active.push_back(Active{std::move(g), std::move(batch), bytes, t_admit});
}

// One tick of a lane: the next level of every group in it. Ticks are
// enqueued without waiting for them, at most kTicksInFlight ahead of the
// device; a group whose last level is enqueued gets its distance gather
// enqueued behind it and moves to completing_.
bool Server::tick(int lane) {
  constexpr int kTicksInFlight = 2;
  Lane& L = lanes_[lane];
  while (!L.ticks.empty() && hipEventQuery(L.ticks.front()) == hipSuccess) {
    free_ev_.push_back(L.ticks.front());
    L.ticks.pop_front();
  }
  (void)hipGetLastError();  // a not-ready query leaves its status behind
  if (L.active.empty() || (int)L.ticks.size() >= kTicksInFlight) return false;
  std::lock_guard<std::mutex> lk(ctx_->mutex());
  // Groups planned under a cost model that has since changed cannot finish:
  // their queries fail, the other groups go on.
  for (size_t i = 0; i < L.active.size();) {
    Active& a = L.active[i];
    if (a.g->cost_epoch == ctx_->cost_epoch()) {
      ++i;
      continue;
    }
    publish(a.reqs, ERR_INVALID_ARG, nullptr, a.t_admit);
    inflight_bytes_ -= a.bytes;
    recycle(std::move(a.g));
    L.active.erase(L.active.begin() + i);
  }
  if (L.active.empty()) return false;
  std::vector<Group*> gs;
  for (auto& a : L.active) gs.push_back(a.g.get());
  const int st = ctx_->launch_levels(gs.data(), (int)gs.size(), nullptr, lane);
  hipEvent_t ev = nullptr;
  if (free_ev_.empty()) {
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
  } else {
    ev = free_ev_.back();
    free_ev_.pop_back();
  }
  if (ev && hipEventRecord(ev, ctx_->lane_stream(lane)) == hipSuccess) L.ticks.push_back(ev);
  else if (ev) free_ev_.push_back(ev);
  for (size_t i = 0; i < L.active.size();) {
    Active& a = L.active[i];
    if (st == OK && a.g->next_level < a.g->nlevels) {
      ++i;
      continue;
    }
    const int fs = st == OK ? ctx_->finish_async(*a.g, lane) : st;
    if (fs != OK) {
      publish(a.reqs, fs, nullptr, a.t_admit);
      inflight_bytes_ -= a.bytes;
      recycle(std::move(a.g));
    } else {
      completing_.push_back(std::move(a));
    }
    L.active.erase(L.active.begin() + i);
  }
  return true;
}

// A group that leaves on an error path may still have launches in flight
// (earlier ticks, or launches of the failing tick made before the failure):
// they finish before the group's buffers are handed to a new plan.
void Server::recycle(std::unique_ptr<Group> g) {
  (void)ctx_->drain();
  spare_.push_back(std::move(g));
}

void Server::loop() {
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      bool busy = !completing_.empty();
      for (const Lane& L : lanes_) busy |= !L.active.empty();
      if (!busy)
        cv_work_.wait(lk, [&]() { return stop_ || !pending_[0].empty() || !pending_[1].empty(); });
      if (stop_ && pending_[0].empty() && pending_[1].empty() && !busy) break;
    }
    bool progressed = false;
    for (int lane = 0; lane < kLanes; ++lane) {
      admit(lane);
      progressed |= tick(lane);
    }
    // Completions; block on the oldest only when nothing else can proceed.
    bool idle = true;
    for (const Lane& L : lanes_) idle &= L.active.empty();
    for (size_t i = 0; i < completing_.size();) {
      Active& a = completing_[i];
      std::vector<int32_t> dist(a.reqs.size(), -1);
      const int ps = ctx_->poll(*a.g, idle && i == 0, dist.data());
      if (ps == ERR_BUSY) {
        ++i;
        continue;
      }
      publish(a.reqs, ps, dist.data(), a.t_admit);
      inflight_bytes_ -= a.bytes;
      spare_.push_back(std::move(a.g));
      completing_.erase(completing_.begin() + i);
      progressed = true;
    }
    // Nothing in flight: memory that grown groups handed over can be freed
    // now without holding up any lane.
    bool empty = completing_.empty();
    for (const Lane& L : lanes_) empty &= L.active.empty();
    if (empty) {
      std::lock_guard<std::mutex> lk(ctx_->mutex());
      if (!ctx_->retire_list()->empty()) (void)ctx_->release_retired();
    }
    // Both lanes two ticks ahead of the device and nothing landed: give the
    // device a moment (new queries still wake the wait).
    if (!progressed) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait_for(lk, std::chrono::microseconds(50));
    }
  }
  for (auto& L : lanes_)
    for (hipEvent_t e : L.ticks) (void)hipEventDestroy(e);
  for (hipEvent_t e : free_ev_) (void)hipEventDestroy(e);
}

}  // namespace xted
