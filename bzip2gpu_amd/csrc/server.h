// server.h - continuous-batching request loop over the batch engine.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace bzg {

class Server {
 public:
  // Device memory one input byte needs inside a pipeline run (workspace
  // plus output), used for admission against the HBM budget.
  static constexpr uint64_t kBytesPerInputByte = 80;

  struct Req {
    const uint8_t* d = nullptr;
    uint64_t len = 0;
    int level = 9;
    int64_t arrival_ns = 0;
    int64_t done_ns = 0;
    int64_t out_bytes = 0;
    int64_t out_off = -1;
    int32_t batch = 0;  // requests in the pipeline run that served it
    int err = 0;
    int state = 0;      // 0 queued, 1 running, 2 done
  };

  // mode 0: continuous batching (batched pipeline); mode 1: plain path,
  // one request per run.  Per-worker batch size is capped so that all
  // workers together stay inside hbm_budget_bytes.
  Server(int device, int mode, int workers, uint64_t max_batch_bytes, uint64_t hbm_budget_bytes);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Queues a request (device buffer); arrival_ns < 0 stamps "now".  Ticket
  // ids increase monotonically over the server's life.
  int64_t submit(const uint8_t* d_in, uint64_t len, int level, int64_t arrival_ns = -1);
  // Blocks until request id is done, copies its record and forgets it: each
  // ticket is waited for once (a second wait, or an unknown id, returns
  // kUnknownTicket).  Otherwise returns the request's error (0 or a HIP code).
  static constexpr int kUnknownTicket = -1;
  int wait(int64_t id, Req* out);
  // Keep every output in a caller-provided device arena (results stay in HBM).
  // Waits until no request is queued or running first.
  void set_keep(bool keep, uint8_t* arena, uint64_t arena_bytes);
  // Waits until no request is queued or running, then restarts the clock
  // (between replays).  Records of finished requests nobody has waited for
  // yet stay, with their times on the old clock.
  void reset_clock();
  int64_t now_ns() const;
  uint64_t max_batch() const { return max_batch_; }
  int64_t batches() const { return batches_; }

 private:
  void worker(int k);
  // arena offset for `bytes` of output, -1 when it does not fit (mu_ held)
  int64_t take_arena_locked(int64_t bytes);
  // waits until nothing is queued or running (mu_ held through lk)
  void quiesce(std::unique_lock<std::mutex>& lk);

  int device_;
  int mode_;
  int nworkers_;
  uint64_t max_batch_;
  uint64_t budget_;
  std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  // every field below is guarded by mu_
  std::unordered_map<int64_t, Req> reqs_;  // requests not yet waited for
  int64_t next_id_ = 0;
  std::deque<int64_t> pending_;
  uint64_t inflight_ = 0;   // device bytes of the running batches
  int running_ = 0;         // batches taken by workers and not yet done
  int64_t batches_ = 0;
  bool stop_ = false;
  bool keep_ = false;
  uint8_t* ext_arena_ = nullptr;
  uint64_t ext_arena_bytes_ = 0;
  int64_t arena_pos_ = 0;
  std::vector<std::thread> threads_;
};

}  // namespace bzg
