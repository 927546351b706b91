// server.hip - request loop with continuous batching (host code).
//
// Requests (device buffers + level) are queued FIFO as they arrive.  A pool
// of pipeline workers, each with its own Engine (workspace) and HIP stream,
// repeatedly takes every queued request that fits its batch budget and
// compresses them together as one batched pipeline run; requests arriving
// while other batches are on the GPU wait for the next free worker, so
// batches of different workers overlap on the device.  Admission keeps the
// device memory that in-flight batches need under the HBM budget.  In plain
// mode every request runs alone through the per-block path instead.
#include "server.h"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace bzg {

namespace {
using Clock = std::chrono::steady_clock;
int64_t ns_since(Clock::time_point t0) {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count();
}
}  // namespace

Server::Server(int device, int mode, int workers, uint64_t max_batch_bytes, uint64_t hbm_budget_bytes)
    : device_(device), mode_(mode), nworkers_(std::max(1, std::min(workers, 8))),
      budget_(hbm_budget_bytes), t0_(Clock::now()) {
  // each worker's workspace grows to about kBytesPerInputByte x its largest batch
  uint64_t per_worker = budget_ / (uint64_t)nworkers_;
  uint64_t cap = per_worker / kBytesPerInputByte;
  max_batch_ = std::max<uint64_t>(1u << 20, std::min<uint64_t>(max_batch_bytes, cap));
  for (int k = 0; k < nworkers_; k++) threads_.emplace_back(&Server::worker, this, k);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto& t : threads_) t.join();
}

int64_t Server::now_ns() const { return ns_since(t0_); }

int64_t Server::submit(const uint8_t* d_in, uint64_t len, int level, int64_t arrival_ns) {
  std::lock_guard<std::mutex> lk(mu_);
  const int64_t id = next_id_++;
  Req& r = reqs_[id];
  r.d = d_in;
  r.len = len;
  r.level = std::min(9, std::max(1, level));
  r.arrival_ns = arrival_ns >= 0 ? arrival_ns : ns_since(t0_);
  pending_.push_back(id);
  cv_work_.notify_one();
  return id;
}

int Server::wait(int64_t id, Req* out) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) return kUnknownTicket;
  // element references of an unordered_map survive inserts and rehashes
  Req& r = it->second;
  cv_done_.wait(lk, [&] { return r.state == 2; });
  if (out) *out = r;
  const int err = r.err;
  reqs_.erase(id);
  return err;
}

void Server::quiesce(std::unique_lock<std::mutex>& lk) {
  cv_done_.wait(lk, [&] { return pending_.empty() && running_ == 0; });
}

void Server::set_keep(bool keep, uint8_t* arena, uint64_t arena_bytes) {
  std::unique_lock<std::mutex> lk(mu_);
  quiesce(lk);
  keep_ = keep;
  ext_arena_ = arena;
  ext_arena_bytes_ = arena_bytes;
  arena_pos_ = 0;
}

void Server::reset_clock() {
  std::unique_lock<std::mutex> lk(mu_);
  quiesce(lk);
  t0_ = Clock::now();
}

void Server::worker(int k) {
  if (hipSetDevice(device_) != hipSuccess) return;
  Engine eng(device_);
  // every admitted batch is a new plan, so capturing stage graphs per batch
  // would not pay off: the workers launch eagerly
  eng.set_graphs(false);
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return;
  std::vector<int64_t> batch;
  std::vector<StreamSpec> ss;
  std::vector<StreamState> lay;
  std::vector<uint8_t> hout;
  for (;;) {
    uint64_t est = 0;
    bool keep = false;        // output placement of this batch, fixed at admission
    uint8_t* arena = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
      if (stop_ && pending_.empty()) break;
      // admission: FIFO, up to the batch cap, within the HBM budget
      batch.clear();
      uint64_t tot = 0;
      while (!pending_.empty()) {
        const Req& r = reqs_.at(pending_.front());
        uint64_t need = (tot + r.len) * kBytesPerInputByte;
        bool fits_batch = batch.empty() || tot + r.len <= max_batch_;
        bool fits_budget = inflight_ + need <= budget_ || (batch.empty() && inflight_ == 0);
        if (mode_ == 1 && !batch.empty()) break;  // plain: one request per run
        if (!fits_batch || !fits_budget) break;
        batch.push_back(pending_.front());
        pending_.pop_front();
        tot += r.len;
      }
      if (batch.empty()) {  // budget exhausted: wait for a completion
        cv_done_.wait(lk, [&] { return stop_ || inflight_ == 0; });
        continue;
      }
      est = tot * kBytesPerInputByte;
      inflight_ += est;
      running_++;
      for (int64_t id : batch) reqs_.at(id).state = 1;
      keep = keep_;
      arena = ext_arena_;
    }
    int err = 0;
    std::vector<std::pair<int64_t, int64_t>> res(batch.size(), {0, -1});  // bytes, arena offset
    if (mode_ == 1) {
      Req r;
      {
        std::lock_guard<std::mutex> lk(mu_);
        r = reqs_.at(batch[0]);
      }
      err = compress_plain_device(eng, r.d, r.len, r.level, 100000 * r.level, hout, s);
      if (!err) {
        res[0].first = (int64_t)hout.size();
        if (keep) {
          int64_t off;
          {
            std::lock_guard<std::mutex> lk(mu_);
            off = take_arena_locked(res[0].first);
          }
          if (off >= 0) {
            err = (int)hipMemcpyAsync(arena + off, hout.data(), hout.size(), hipMemcpyHostToDevice, s);
            if (!err) err = (int)hipStreamSynchronize(s);
            res[0].second = off;
          }
        }
      }
    } else {
      ss.clear();
      const uint8_t* base = nullptr;
      {
        std::lock_guard<std::mutex> lk(mu_);
        // one device address space: offsets relative to the lowest buffer
        base = reqs_.at(batch[0]).d;
        for (int64_t id : batch) base = std::min(base, reqs_.at(id).d);
        for (int64_t id : batch) {
          const Req& r = reqs_.at(id);
          ss.push_back(StreamSpec{(uint64_t)(r.d - base), r.len, r.level, 100000 * r.level});
        }
      }
      if (!err) err = eng.plan(ss);
      if (!err) err = eng.setup(base, s);
      if (!err) err = eng.run(s, true, nullptr);
      if (!err) err = eng.stream_layout(lay, s);
      if (!err) {
        for (size_t q = 0; q < batch.size(); q++) {
          res[q].first = lay[q].bytes;
          if (keep) {
            int64_t off;
            {
              std::lock_guard<std::mutex> lk(mu_);
              off = take_arena_locked(lay[q].bytes);
            }
            if (off >= 0) {
              hipError_t he = hipMemcpyAsync(arena + off, eng.out_buffer() + lay[q].byte_off,
                                             (size_t)lay[q].bytes, hipMemcpyDeviceToDevice, s);
              if (he != hipSuccess) err = (int)he;
              res[q].second = off;
            }
          }
        }
        if (!err && keep) err = (int)hipStreamSynchronize(s);
      }
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      int64_t t = ns_since(t0_);
      for (size_t q = 0; q < batch.size(); q++) {
        Req& r = reqs_.at(batch[q]);
        r.done_ns = t;
        r.out_bytes = res[q].first;
        r.out_off = res[q].second;
        r.err = err;
        r.batch = (int32_t)batch.size();
        r.state = 2;
      }
      inflight_ -= est;
      running_--;
      batches_++;
    }
    cv_done_.notify_all();
    cv_work_.notify_all();
  }
  (void)hipStreamDestroy(s);
}

int64_t Server::take_arena_locked(int64_t bytes) {
  int64_t b = (bytes + 3) & ~3ll;
  uint64_t cap = ext_arena_ ? ext_arena_bytes_ : 0;
  if ((uint64_t)(arena_pos_ + b) > cap) return -1;
  int64_t off = arena_pos_;
  arena_pos_ += b;
  return off;
}

}  // namespace bzg
