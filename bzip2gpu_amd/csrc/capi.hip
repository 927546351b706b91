// capi.hip - the extern "C" boundary (include/bzg.h) over the batch engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bzg.h"
#include "decode.h"
#include "engine.h"
#include "server.h"
#include "tuner.h"

using namespace bzg;

namespace {

thread_local std::string g_err;
thread_local int g_dev = 0;

// Workspace of the allocation-free entry point
// (bzg_compress_batch_device_noalloc): a fixed array of this library in
// device memory, one copy per GPU the library runs on, held until the
// process ends.  In host code `scratch` only names the array; its device
// address comes from hipGetSymbolAddress, per device (Ctx::scratch_ptr).
constexpr size_t kScratchBytes = size_t(BZG_NOALLOC_SCRATCH_MIB) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

struct Ctx {
  explicit Ctx(int dev) : eng(dev), stage(dev), noalloc(dev), dec(dev), device(dev) {}
  Engine eng;      // batched / plain paths
  Engine stage;    // per-stage entry points
  Engine noalloc;  // the allocation-free batched path, in `scratch`
  // engines and streams of the reference-style threaded plain path
  std::vector<std::unique_ptr<Engine>> pool;
  std::vector<hipStream_t> pool_streams;
  void* scratch_ptr = nullptr;  // this device's copy of `scratch`
  int scratch_err = 0;          // set-up failure of the allocation-free path
  Decoder dec;     // .bz2 -> raw
  DecodeTimes dlast{};
  int device;
  std::mutex mu;
  uint8_t* d_in = nullptr;
  size_t d_in_cap = 0;
  hipStream_t stream = nullptr;
  StageTimes last{};
  int64_t last_totals[kTotals] = {};  // of the last batched call
  uint64_t max_batch_bytes = 1ull << 30;  // raw bytes per device sub-batch
  uint8_t* d_res = nullptr;                // multi-sub-batch results
  size_t d_res_cap = 0;
  // second engine and stream of overlapped sub-batches
  std::unique_ptr<Engine> eng2;
  hipStream_t stream2 = nullptr;
  bool overlap = true;
  ~Ctx() {
    (void)hipSetDevice(device);
    for (auto ps : pool_streams) (void)hipStreamDestroy(ps);
    if (stream2) (void)hipStreamDestroy(stream2);
  }
};

std::mutex g_mu;
std::unique_ptr<Ctx> g_ctx[64];

bzg_status fail(bzg_status code, const std::string& msg) {
  g_err = msg;
  return code;
}

bzg_status hip_fail(int e, const char* where) {
  if (e == (int)hipErrorIllegalState)
    return fail(BZG_ERR_INTERNAL, std::string(where) + ": device-side invariant failed");
  return fail(BZG_ERR_HIP, std::string(where) + ": " + hipGetErrorString((hipError_t)e));
}

Ctx* ctx_for(int dev, bzg_status* st) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    *st = fail(BZG_ERR_HIP, "no HIP device available");
    return nullptr;
  }
  if (dev < 0 || dev >= n || dev >= 64) {
    *st = fail(BZG_ERR_INVALID, "device index out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx[dev]) {
    if (hipSetDevice(dev) != hipSuccess) {
      *st = fail(BZG_ERR_HIP, "hipSetDevice failed");
      return nullptr;
    }
    g_ctx[dev].reset(new Ctx(dev));
    Ctx* c = g_ctx[dev].get();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      g_ctx[dev].reset();
      *st = fail(BZG_ERR_HIP, "hipStreamCreate failed");
      return nullptr;
    }
    // Set-up of the allocation-free path, once per device with the device
    // current: the device address of this GPU's `scratch`, the engine that
    // runs in it (eager launches, round events made now).
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) {
      c->scratch_ptr = p;
      c->noalloc.use_fixed_workspace(p, kScratchBytes);
      c->noalloc.set_graphs(false);
      c->scratch_err = c->noalloc.reserve_round_events(kMaxRounds);
    } else {
      c->scratch_err = e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
    }
  }
  *st = BZG_OK;
  return g_ctx[dev].get();
}

int ensure_input(Ctx* c, size_t bytes) {
  if (bytes <= c->d_in_cap) return 0;
  if (c->d_in) {
    hipError_t e = hipFree(c->d_in);
    if (e != hipSuccess) return (int)e;
  }
  c->d_in = nullptr;
  c->d_in_cap = 0;
  size_t want = bytes + bytes / 8 + 256;
  hipError_t e = hipMalloc(&c->d_in, want);
  if (e != hipSuccess) return (int)e;
  c->d_in_cap = want;
  return 0;
}

// Grow-only result buffer of the multi-sub-batch path; its contents need
// not survive a growth (a call sizes it before any sub-batch runs).
int ensure_result(Ctx* c, size_t bytes) {
  if (bytes <= c->d_res_cap) return 0;
  if (c->d_res) {
    hipError_t e = hipFree(c->d_res);
    if (e != hipSuccess) return (int)e;
  }
  c->d_res = nullptr;
  c->d_res_cap = 0;
  const size_t want = bytes + bytes / 8 + 4096;
  hipError_t e = hipMalloc(&c->d_res, want);
  if (e != hipSuccess) return (int)e;
  c->d_res_cap = want;
  return 0;
}

int clamp_level(int level) { return level < 1 ? 1 : (level > 9 ? 9 : level); }

}  // namespace

extern "C" {

const char* bzg_version(void) { return "bzip2gpu_amd 0.1 (gfx950)"; }

const char* bzg_last_error(void) { return g_err.c_str(); }

bzg_status bzg_init(int device) {
  bzg_status st;
  Ctx* c = ctx_for(device, &st);
  if (!c) return st;
  g_dev = device;
  return BZG_OK;
}

size_t bzg_compress_bound(size_t len, int32_t block_bytes) {
  if (block_bytes <= 0 || block_bytes > BZG_REF_BLOCK_BYTES) block_bytes = BZG_REF_BLOCK_BYTES;
  size_t nb = (len + (size_t)block_bytes - 1) / (size_t)block_bytes;
  int32_t cap = cap_for(block_bytes);
  size_t per = (size_t)(hdr_words_for(cap) + pay_words_for(cap)) * 4;
  return nb * per + 64;
}

bzg_status bzg_compress_buffer(const uint8_t* h_in, size_t len, int level, int32_t block_bytes,
                               uint8_t* h_out, size_t out_cap, size_t* out_len) {
  if ((!h_in && len) || !h_out || !out_len) return fail(BZG_ERR_INVALID, "null pointer");
  if (block_bytes <= 0 || block_bytes > BZG_REF_BLOCK_BYTES) return fail(BZG_ERR_INVALID, "block_bytes out of range");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  int e = ensure_input(c, len + 16);
  if (e) return hip_fail(e, "input buffer");
  if (len) {
    hipError_t he = hipMemcpyAsync(c->d_in, h_in, len, hipMemcpyHostToDevice, c->stream);
    if (he != hipSuccess) return hip_fail((int)he, "H2D");
  }
  std::vector<StreamSpec> ss{StreamSpec{0, len, clamp_level(level), block_bytes}};
  if ((e = c->eng.plan(ss))) return hip_fail(e, "plan");
  if ((e = c->eng.setup(c->d_in, c->stream))) return hip_fail(e, "setup");
  if ((e = c->eng.run(c->stream, true, &c->last))) return hip_fail(e, "run");
  std::vector<StreamState> lay;
  if ((e = c->eng.stream_layout(lay, c->stream))) return hip_fail(e, "layout");
  size_t n = (size_t)lay[0].bytes;
  *out_len = n;
  if (n > out_cap) return fail(BZG_ERR_CAPACITY, "output buffer too small");
  hipError_t he = hipMemcpyAsync(h_out, c->eng.out_buffer() + lay[0].byte_off, n, hipMemcpyDeviceToHost, c->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  if (he != hipSuccess) return hip_fail((int)he, "D2H");
  return BZG_OK;
}

namespace {

// Shared body of the plain path: d_in is either the caller's device buffer
// or the context's staging buffer refilled per block from h_in.
bzg_status plain_run(Ctx* c, const uint8_t* h_in, const uint8_t* d_in, size_t len, int level, int32_t block_bytes,
                     uint8_t* h_out, size_t out_cap, size_t* out_len, hipStream_t s) {
  std::vector<uint8_t> out;
  int e;
  if (d_in) {
    if ((e = compress_plain_device(c->eng, d_in, len, level, block_bytes, out, s))) return hip_fail(e, "plain");
  } else {
    // host input: one copy into the context's staging buffer, then the same
    // per-block device path
    if ((e = ensure_input(c, len + 16))) return hip_fail(e, "input buffer");
    if (len && (e = (int)hipMemcpyAsync(c->d_in, h_in, len, hipMemcpyHostToDevice, s))) return hip_fail(e, "H2D");
    if ((e = compress_plain_device(c->eng, c->d_in, len, level, block_bytes, out, s))) return hip_fail(e, "plain");
  }
  *out_len = out.size();
  if (out.size() > out_cap) return fail(BZG_ERR_CAPACITY, "output buffer too small");
  memcpy(h_out, out.data(), out.size());
  return BZG_OK;
}

}  // namespace

bzg_status bzg_compress_plain(const uint8_t* h_in, size_t len, int level, int32_t block_bytes,
                              uint8_t* h_out, size_t out_cap, size_t* out_len) {
  if ((!h_in && len) || !h_out || !out_len) return fail(BZG_ERR_INVALID, "null pointer");
  if (block_bytes <= 0 || block_bytes > BZG_REF_BLOCK_BYTES) return fail(BZG_ERR_INVALID, "block_bytes out of range");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  return plain_run(c, h_in ? h_in : reinterpret_cast<const uint8_t*>(""), nullptr, len, level, block_bytes, h_out,
                   out_cap, out_len, c->stream);
}

bzg_status bzg_compress_plain_device(const uint8_t* d_in, size_t len, int level, int32_t block_bytes,
                                     uint8_t* h_out, size_t out_cap, size_t* out_len, void* stream) {
  if ((!d_in && len) || !h_out || !out_len) return fail(BZG_ERR_INVALID, "null pointer");
  if (block_bytes <= 0 || block_bytes > BZG_REF_BLOCK_BYTES) return fail(BZG_ERR_INVALID, "block_bytes out of range");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return plain_run(c, nullptr, d_in, len, level, block_bytes, h_out, out_cap, out_len, s);
}

bzg_status bzg_compress_batch_device(const uint8_t* d_in, int32_t n_inputs, const uint64_t* in_off,
                                     const uint64_t* lens, const int32_t* levels, int32_t block_bytes,
                                     uint8_t** d_out, int64_t* h_byte_off, int64_t* h_bytes, void* stream) {
  if (n_inputs < 0 || (n_inputs > 0 && (!d_in || !in_off || !lens || !levels)) || !d_out || !h_byte_off || !h_bytes)
    return fail(BZG_ERR_INVALID, "null pointer");
  if (block_bytes > BZG_REF_BLOCK_BYTES) return fail(BZG_ERR_INVALID, "block_bytes out of range");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  // Inputs are taken in order in sub-batches of at most max_batch_bytes raw
  // bytes (bounded workspace).  With more than one sub-batch, each one
  // assembles its streams directly into its own region of the
  // library-owned result buffer (sized from every sub-batch's plan first),
  // so no stream is copied afterwards.
  const uint64_t max_batch = c->max_batch_bytes;
  std::vector<std::pair<int, int>> parts;  // [i0, i1) per sub-batch
  for (int i0 = 0; i0 < n_inputs;) {
    uint64_t tot = 0;
    int i1 = i0;
    while (i1 < n_inputs && (i1 == i0 || tot + lens[i1] <= max_batch)) tot += lens[i1++];
    parts.emplace_back(i0, i1);
    i0 = i1;
  }
  auto specs = [&](int i0, int i1, std::vector<StreamSpec>& ss) {
    ss.clear();
    for (int i = i0; i < i1; i++) {
      const int lv = clamp_level(levels[i]);
      ss.push_back(StreamSpec{in_off[i], lens[i], lv, block_bytes > 0 ? block_bytes : 100000 * lv});
    }
  };
  std::vector<StreamSpec> ss;
  std::vector<int64_t> region(parts.size() + 1, 0);  // result-buffer offset of each sub-batch
  int e;
  bool tuning = false;  // some sub-batch would time autotuner variants
  if (parts.size() > 1) {
    for (size_t p = 0; p < parts.size(); p++) {
      specs(parts[p].first, parts[p].second, ss);
      if ((e = c->eng.plan(ss))) return hip_fail(e, "plan");
      region[p + 1] = region[p] + ((c->eng.planned_out_bytes() + 255) & ~255ll);
      tuning = tuning || c->eng.will_tune();
    }
    if ((e = ensure_result(c, (size_t)region.back()))) return hip_fail(e, "result buffer");
  }
  // Sub-batch p runs on engine p % 2 (its own HIP stream and host thread
  // when there are two or more sub-batches and overlap is on); the front
  // ends and suffix sorts run one at a time in sub-batch order (a gate the
  // engines take before their front end and leave when their sort has
  // ended), so sub-batch p's MTF, Huffman, pack and assembly overlap
  // sub-batch p+1's front end and sort.
  // (runs that time autotuner variants take the sub-batches one at a time,
  // so that the other pipeline does not distort the timings)
  const bool two = parts.size() > 1 && c->overlap && !tuning;
  if (two && !c->eng2) {
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess)
      return fail(BZG_ERR_HIP, "hipStreamCreate");
    c->eng2.reset(new Engine(c->device));
  }
  Engine* engs[2] = {&c->eng, two ? c->eng2.get() : &c->eng};
  hipStream_t strs[2] = {s, two ? c->stream2 : s};
  std::vector<StageTimes> ptimes(parts.size());
  std::vector<std::array<int64_t, kTotals>> ptot(parts.size());
  struct Gate {
    std::mutex m;
    std::condition_variable cv;
    size_t next = 0;
    bool abort = false;
  } gate;
  std::atomic<int> first_err{0};
  std::string err_where;
  std::mutex err_mu;
  auto note = [&](int e, const char* where) {
    std::lock_guard<std::mutex> lk(err_mu);
    if (!first_err.load()) {
      first_err = e;
      err_where = where;
    }
    std::lock_guard<std::mutex> lg(gate.m);
    gate.abort = true;
    gate.cv.notify_all();
  };
  auto worker = [&](int j) {
    if (hipSetDevice(c->device) != hipSuccess) return note((int)hipErrorInvalidDevice, "device");
    Engine& E = *engs[j];
    std::vector<StreamSpec> wss;
    std::vector<StreamState> wlay;
    for (size_t p = (size_t)j; p < parts.size(); p += two ? 2 : 1) {
      if (two)
        E.set_sort_gate([&gate, p](bool enter) -> int {
          std::unique_lock<std::mutex> lk(gate.m);
          if (enter) {
            gate.cv.wait(lk, [&] { return gate.next == p || gate.abort; });
            return gate.abort ? (int)hipErrorNotReady : 0;  // the other engine failed
          }
          gate.next = p + 1;
          gate.cv.notify_all();
          return 0;
        });
      const int i0 = parts[p].first, i1 = parts[p].second;
      specs(i0, i1, wss);
      uint8_t* dst = parts.size() > 1 ? c->d_res + region[p] : nullptr;
      int e2;
      if ((e2 = E.plan(wss))) return note(e2, "plan");
      if ((e2 = E.setup(d_in, strs[j], dst))) return note(e2, "setup");
      if ((e2 = E.run(strs[j], true, &ptimes[p]))) return note(e2, "run");
      if ((e2 = E.stream_layout(wlay, strs[j]))) return note(e2, "layout");
      for (int k = 0; k < kTotals; k++) ptot[p][k] = E.totals()[k];
      for (int i = i0; i < i1; i++) {
        h_byte_off[i] = region[p] + wlay[i - i0].byte_off;
        h_bytes[i] = wlay[i - i0].bytes;
      }
    }
  };
  if (two) {
    // the second stream starts after the caller's stream's earlier work
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail(BZG_ERR_HIP, "event");
    if (hipEventRecord(ev, s) != hipSuccess || hipStreamWaitEvent(c->stream2, ev, 0) != hipSuccess) {
      (void)hipEventDestroy(ev);
      return fail(BZG_ERR_HIP, "stream order");
    }
    std::thread t1(worker, 1);
    worker(0);
    t1.join();
    for (Engine* E : engs) E->set_sort_gate(nullptr);
    // the caller's stream continues after the second stream's work
    hipError_t he2 = hipEventRecord(ev, c->stream2);
    if (he2 == hipSuccess) he2 = hipStreamWaitEvent(s, ev, 0);
    (void)hipEventDestroy(ev);
    if (!first_err.load() && he2 != hipSuccess) note((int)he2, "stream order");
  } else {
    worker(0);
  }
  if (const int e3 = first_err.load()) return hip_fail(e3, err_where.c_str());
  StageTimes sum{};
  int64_t totals[kTotals] = {};
  for (size_t p = 0; p < parts.size(); p++) {
    const StageTimes& t = ptimes[p];
    sum.front += t.front; sum.bwt += t.bwt; sum.mtf += t.mtf; sum.huff += t.huff;
    sum.pack += t.pack; sum.assemble += t.assemble; sum.rounds = std::max(sum.rounds, t.rounds);
    sum.bwt_elems += t.bwt_elems;
    sum.small_rows += t.small_rows;
    sum.large_rows += t.large_rows;
    sum.large_pass_rows += t.large_pass_rows;
    for (int k = 0; k < kTotals; k++) totals[k] += ptot[p][k];
  }
  *d_out = parts.size() > 1 ? c->d_res : c->eng.out_buffer();
  hipError_t he = hipStreamSynchronize(s);
  if (he != hipSuccess) return hip_fail((int)he, "batch");
  c->last = sum;
  std::memcpy(c->last_totals, totals, sizeof totals);
  return BZG_OK;
}

bzg_status bzg_last_batch_totals(int64_t* h_totals6) {
  if (!h_totals6) return fail(BZG_ERR_INVALID, "null pointer");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  for (int k = 0; k < kTotals; k++) h_totals6[k] = c->last_totals[k];
  return BZG_OK;
}

bzg_status bzg_compress_plain_threads(const uint8_t* d_in, int32_t n_inputs, const uint64_t* in_off,
                                      const uint64_t* lens, const int32_t* levels, int32_t block_bytes,
                                      int32_t threads, uint8_t* h_out, size_t out_cap, int64_t* h_byte_off,
                                      int64_t* h_bytes) {
  if (n_inputs < 0 || (n_inputs > 0 && (!d_in || !in_off || !lens || !levels)) || !h_out || !h_byte_off ||
      !h_bytes)
    return fail(BZG_ERR_INVALID, "null pointer");
  if (block_bytes > BZG_REF_BLOCK_BYTES) return fail(BZG_ERR_INVALID, "block_bytes out of range");
  if (threads < 1 || threads > 16) return fail(BZG_ERR_INVALID, "threads must be 1..16");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  while ((int)c->pool.size() < threads) {
    hipStream_t ps = nullptr;
    if (hipStreamCreateWithFlags(&ps, hipStreamNonBlocking) != hipSuccess) return fail(BZG_ERR_HIP, "hipStreamCreate");
    c->pool.emplace_back(new Engine(c->device));
    c->pool.back()->set_graphs(false);
    c->pool_streams.push_back(ps);
  }
  // the work list: every block of every input, in input and block order
  struct Job { int input; uint64_t off; int32_t n, level; };
  std::vector<Job> jobs;
  std::vector<size_t> first(n_inputs + 1, 0);
  for (int i = 0; i < n_inputs; i++) {
    const int lv = clamp_level(levels[i]);
    const uint64_t bb = block_bytes > 0 ? (uint64_t)block_bytes : 100000ull * (uint64_t)lv;
    first[i] = jobs.size();
    for (uint64_t o = 0; o < lens[i]; o += bb) jobs.push_back(Job{i, in_off[i] + o, (int32_t)std::min(bb, lens[i] - o), lv});
  }
  first[n_inputs] = jobs.size();
  std::vector<BlockResult> res(jobs.size());
  std::atomic<size_t> next{0};
  std::atomic<int> err{0};
  auto work = [&](int t) {
    if (hipSetDevice(c->device) != hipSuccess) { err = (int)hipErrorInvalidDevice; return; }
    for (size_t j; !err && (j = next++) < jobs.size();) {
      const int e = compress_one_block(*c->pool[t], d_in + jobs[j].off, jobs[j].n, jobs[j].level, res[j],
                                       c->pool_streams[t]);
      if (e) err = e;
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < threads; t++) th.emplace_back(work, t);
  for (auto& x : th) x.join();
  if (err) return hip_fail(err, "plain threads");
  size_t pos = 0;
  std::vector<uint8_t> one;
  for (int i = 0; i < n_inputs; i++) {
    one.clear();
    serialize_stream(levels[i], res.data() + first[i], first[i + 1] - first[i], one);
    if (pos + one.size() > out_cap) return fail(BZG_ERR_CAPACITY, "output buffer too small");
    memcpy(h_out + pos, one.data(), one.size());
    h_byte_off[i] = (int64_t)pos;
    h_bytes[i] = (int64_t)one.size();
    pos += one.size();
  }
  return BZG_OK;
}

bzg_status bzg_compress_batch_device_noalloc(const uint8_t* d_in, int32_t n_inputs, const uint64_t* in_off,
                                             const uint64_t* lens, const int32_t* levels, int32_t block_bytes,
                                             uint8_t* d_out, size_t out_cap, int64_t* h_byte_off, int64_t* h_bytes,
                                             void* stream) {
  if (n_inputs < 0 || (n_inputs > 0 && (!d_in || !in_off || !lens || !levels || !d_out)) || !h_byte_off || !h_bytes)
    return fail(BZG_ERR_INVALID, "null pointer");
  if (block_bytes > BZG_REF_BLOCK_BYTES) return fail(BZG_ERR_INVALID, "block_bytes out of range");
  uint64_t total = 0, blocks = 0;
  for (int i = 0; i < n_inputs; i++) {
    const uint64_t bb = block_bytes > 0 ? (uint64_t)block_bytes : 100000ull * (uint64_t)clamp_level(levels[i]);
    total += lens[i];
    blocks += (lens[i] + bb - 1) / bb;
  }
  if (total > BZG_NOALLOC_MAX_BYTES || blocks > BZG_NOALLOC_MAX_BLOCKS)
    return fail(BZG_ERR_INVALID, "allocation-free path takes at most " + std::to_string(BZG_NOALLOC_MAX_BYTES) +
                                     " raw bytes in " + std::to_string(BZG_NOALLOC_MAX_BLOCKS) + " blocks (got " +
                                     std::to_string(total) + " bytes in " + std::to_string(blocks) + " blocks)");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->scratch_err) return hip_fail(c->scratch_err, "allocation-free workspace set-up");
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  std::vector<StreamSpec> ss;
  for (int i = 0; i < n_inputs; i++) {
    const int lv = clamp_level(levels[i]);
    ss.push_back(StreamSpec{in_off[i], lens[i], lv, block_bytes > 0 ? block_bytes : 100000 * lv});
  }
  Engine& E = c->noalloc;
  int e;
  if ((e = E.plan(ss))) return hip_fail(e, "plan");
  if ((size_t)E.planned_out_bytes() > out_cap)
    return fail(BZG_ERR_CAPACITY, "output buffer too small: needs " + std::to_string(E.planned_out_bytes()) + " bytes");
  if ((e = E.setup(d_in, s, d_out))) {
    if (e == (int)hipErrorOutOfMemory)
      return fail(BZG_ERR_INTERNAL, "batch needs " + std::to_string(E.needed_workspace()) +
                                        " bytes of workspace, the fixed array holds " + std::to_string(kScratchBytes));
    return hip_fail(e, "setup");
  }
  if ((e = E.run(s, true, nullptr))) return hip_fail(e, "run");
  std::vector<StreamState> lay;
  if ((e = E.stream_layout(lay, s))) return hip_fail(e, "layout");
  for (int i = 0; i < n_inputs; i++) {
    h_byte_off[i] = lay[(size_t)i].byte_off;
    h_bytes[i] = lay[(size_t)i].bytes;
  }
  return BZG_OK;
}

bzg_status bzg_compress_segment(const uint8_t* d_in, size_t len, int32_t block_bytes, int32_t lead_bits,
                                uint32_t* h_crc, uint8_t* h_out, size_t out_cap, size_t* out_len, int64_t* h_bits,
                                void* stream) {
  if ((!d_in && len) || !h_crc || !h_out || !out_len || !h_bits) return fail(BZG_ERR_INVALID, "null pointer");
  if (block_bytes <= 0 || block_bytes > BZG_REF_BLOCK_BYTES) return fail(BZG_ERR_INVALID, "block_bytes out of range");
  if (lead_bits < 0 || lead_bits > 7) return fail(BZG_ERR_INVALID, "lead_bits must be 0..7");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  StreamSpec spec{0, len, 9, block_bytes};
  spec.lead = lead_bits;
  spec.crc_in = *h_crc;
  std::vector<StreamSpec> ss{spec};
  std::vector<StreamState> lay;
  int e;
  if ((e = c->eng.plan(ss))) return hip_fail(e, "plan");
  if ((e = c->eng.setup(d_in, s))) return hip_fail(e, "setup");
  if ((e = c->eng.run(s, true, &c->last))) return hip_fail(e, "run");
  if ((e = c->eng.stream_layout(lay, s))) return hip_fail(e, "layout");
  const size_t n = (size_t)lay[0].bytes;
  *out_len = n;
  if (n > out_cap) return fail(BZG_ERR_CAPACITY, "output buffer too small");
  if (n) {
    hipError_t he = hipMemcpyAsync(h_out, c->eng.out_buffer() + lay[0].byte_off, n, hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) return hip_fail((int)he, "D2H");
  }
  *h_crc = lay[0].crc;
  *h_bits = lay[0].bits;
  return BZG_OK;
}

int32_t bzg_last_round_stats(int64_t* h_rows, float* h_ms, int32_t max_rounds) {
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return -1;
  const auto& rr = c->eng.round_rows();
  const auto& rm = c->eng.round_ms();
  int32_t n = (int32_t)std::min(rr.size(), rm.size());
  for (int32_t i = 0; i < n && i < max_rounds; i++) {
    if (h_rows) h_rows[i] = rr[i];
    if (h_ms) h_ms[i] = rm[i];
  }
  return n;
}

bzg_status bzg_set_max_batch_bytes(uint64_t bytes) {
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  c->max_batch_bytes = bytes < (1u << 20) ? (1u << 20) : bytes;
  return BZG_OK;
}

bzg_status bzg_set_batch_overlap(int32_t on) {
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  c->overlap = on != 0;
  return BZG_OK;
}

struct bzg_server {
  std::unique_ptr<bzg::Server> s;
};

bzg_status bzg_server_create(int32_t mode, int32_t workers, uint64_t max_batch_bytes, uint64_t hbm_budget_bytes,
                             bzg_server** out) {
  if (!out || (mode != 0 && mode != 1) || workers < 1) return fail(BZG_ERR_INVALID, "bad argument");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  bzg_server* srv = new bzg_server;
  srv->s.reset(new bzg::Server(c->device, mode, workers, max_batch_bytes, hbm_budget_bytes));
  *out = srv;
  return BZG_OK;
}

void bzg_server_destroy(bzg_server* srv) { delete srv; }

bzg_status bzg_server_submit(bzg_server* srv, const uint8_t* d_in, uint64_t len, int32_t level, int64_t* h_ticket) {
  if (!srv || !h_ticket || (len && !d_in)) return fail(BZG_ERR_INVALID, "bad argument");
  *h_ticket = srv->s->submit(d_in, len, level, -1);
  return BZG_OK;
}

bzg_status bzg_server_wait(bzg_server* srv, int64_t ticket, int64_t* h_out_bytes, int64_t* h_out_off,
                           int64_t* h_latency_ns) {
  if (!srv) return fail(BZG_ERR_INVALID, "bad argument");
  bzg::Server::Req r;
  int e = srv->s->wait(ticket, &r);
  if (e == bzg::Server::kUnknownTicket) return fail(BZG_ERR_INVALID, "unknown or already waited-for ticket");
  if (e) return hip_fail(e, "server request");
  if (h_out_bytes) *h_out_bytes = r.out_bytes;
  if (h_out_off) *h_out_off = r.out_off;
  if (h_latency_ns) *h_latency_ns = r.done_ns - r.arrival_ns;
  return BZG_OK;
}

bzg_status bzg_server_keep(bzg_server* srv, uint8_t* d_arena, uint64_t arena_bytes) {
  if (!srv) return fail(BZG_ERR_INVALID, "bad argument");
  srv->s->set_keep(d_arena != nullptr, d_arena, arena_bytes);
  return BZG_OK;
}

bzg_status bzg_server_replay(bzg_server* srv, const uint8_t* d_pool, int32_t n, const uint64_t* in_off,
                             const uint64_t* lens, const int32_t* levels, const int64_t* arrival_ns,
                             int64_t* h_done_ns, int64_t* h_out_bytes, int64_t* h_out_off, int32_t* h_batch) {
  if (!srv || n < 0 || (n && (!d_pool || !in_off || !lens || !levels || !arrival_ns || !h_done_ns)))
    return fail(BZG_ERR_INVALID, "bad argument");
  bzg::Server& S = *srv->s;
  S.reset_clock();
  std::vector<int64_t> ids((size_t)n);
  for (int32_t i = 0; i < n; i++) {
    // open loop: submit at the scheduled arrival time, whatever the backlog
    int64_t t = arrival_ns[i];
    for (;;) {
      int64_t now = S.now_ns();
      if (now >= t) break;
      int64_t dt = t - now;
      if (dt > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(dt - 100000));
    }
    ids[i] = S.submit(d_pool + in_off[i], lens[i], levels[i], t);
  }
  for (int32_t i = 0; i < n; i++) {
    bzg::Server::Req r;
    int e = S.wait(ids[i], &r);
    if (e == bzg::Server::kUnknownTicket) return fail(BZG_ERR_INTERNAL, "replay request record lost");
    if (e) return hip_fail(e, "server request");
    h_done_ns[i] = r.done_ns;
    if (h_out_bytes) h_out_bytes[i] = r.out_bytes;
    if (h_out_off) h_out_off[i] = r.out_off;
    if (h_batch) h_batch[i] = r.batch;
  }
  return BZG_OK;
}

bzg_status bzg_probe_kernel(const char* kernel_name) {
  probe_set(kernel_name);
  return BZG_OK;
}

bzg_status bzg_probe_read(float* h_total_ms, int32_t* h_launches) {
  if (!h_total_ms || !h_launches) return fail(BZG_ERR_INVALID, "null pointer");
  int n = 0;
  int e = probe_read(h_total_ms, &n);
  if (e) return hip_fail(e, "probe");
  *h_launches = n;
  return BZG_OK;
}

bzg_status bzg_last_stage_times(float* h_ms6, int32_t* h_rounds, int64_t* h_bwt_rows) {
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  if (h_ms6) {
    h_ms6[0] = c->last.front; h_ms6[1] = c->last.bwt; h_ms6[2] = c->last.mtf;
    h_ms6[3] = c->last.huff; h_ms6[4] = c->last.pack; h_ms6[5] = c->last.assemble;
  }
  if (h_rounds) *h_rounds = c->last.rounds;
  if (h_bwt_rows) *h_bwt_rows = c->last.bwt_elems;
  return BZG_OK;
}

bzg_status bzg_set_stage_graphs(int32_t on) {
  bzg::set_stage_graphs(on != 0);
  return BZG_OK;
}

int64_t bzg_graph_replays(void) {
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return -1;
  int64_t n = 0;
  for (int i = 0; i < Engine::kStageGroups; i++) n += c->eng.graph_replays(i);
  return n;
}

int64_t bzg_graph_captures(void) {
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return -1;
  int64_t n = 0;
  for (int i = 0; i < Engine::kStageGroups; i++) n += c->eng.graph_captures(i);
  return n;
}

bzg_status bzg_tune_force(int32_t knob, int32_t variant) {
  if (knob < 0 || knob >= bzg::TUNE_KNOBS || variant < -1 || variant >= bzg::Tuner::get().variants(knob))
    return fail(BZG_ERR_INVALID, "bad tuning knob or variant");
  bzg::Tuner::get().force(knob, variant);
  return BZG_OK;
}

int32_t bzg_tune_state(int32_t knob, int32_t size_class, int32_t* h_values, float* h_ns_per_unit,
                       int32_t* h_samples, int32_t max_variants) {
  if (knob < 0 || knob >= bzg::TUNE_KNOBS || size_class < 0 || size_class >= bzg::kTuneClasses) return -2;
  bzg::Tuner& t = bzg::Tuner::get();
  double ns[bzg::kTuneMaxVariants];
  int n[bzg::kTuneMaxVariants];
  const int chosen = t.state(knob, size_class, ns, n);
  for (int v = 0; v < t.variants(knob) && v < max_variants; v++) {
    if (h_values) h_values[v] = t.value(knob, v);
    if (h_ns_per_unit) h_ns_per_unit[v] = (float)ns[v];
    if (h_samples) h_samples[v] = n[v];
  }
  return chosen;
}

bzg_status bzg_last_sort_rows(int64_t* h_small_rows, int64_t* h_large_rows, int64_t* h_large_pass_rows) {
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  if (h_small_rows) *h_small_rows = c->last.small_rows;
  if (h_large_rows) *h_large_rows = c->last.large_rows;
  if (h_large_pass_rows) *h_large_pass_rows = c->last.large_pass_rows;
  return BZG_OK;
}

// ------------------------------------------------------------ decompression

namespace {

bzg_status decode_status(Ctx* c, int e) {
  if (e == kDecDataError) return fail(BZG_ERR_DATA, c->dec.error());
  if (e) return hip_fail(e, "decompress");
  return BZG_OK;
}

bzg_status copy_decoded(Ctx* c, uint8_t* h_out, size_t out_cap, size_t* out_len) {
  const size_t n = c->dec.out_len();
  *out_len = n;
  if (n > out_cap) return fail(BZG_ERR_CAPACITY, "output buffer too small");
  if (n) {
    hipError_t he = hipMemcpyAsync(h_out, c->dec.out(), n, hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) return hip_fail((int)he, "D2H");
  }
  return BZG_OK;
}

}  // namespace

bzg_status bzg_decompress_buffer(const uint8_t* h_in, size_t len, uint8_t* h_out, size_t out_cap,
                                 size_t* out_len) {
  if ((!h_in && len) || (!h_out && out_cap) || !out_len) return fail(BZG_ERR_INVALID, "null pointer");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  int e = ensure_input(c, len + 16);
  if (e) return hip_fail(e, "input buffer");
  if (len) {
    hipError_t he = hipMemcpyAsync(c->d_in, h_in, len, hipMemcpyHostToDevice, c->stream);
    if (he != hipSuccess) return hip_fail((int)he, "H2D");
  }
  if ((st = decode_status(c, c->dec.run(c->d_in, len, c->stream, &c->dlast))) != BZG_OK) return st;
  return copy_decoded(c, h_out, out_cap, out_len);
}

bzg_status bzg_decompress_fetch(uint8_t* h_out, size_t out_cap, size_t* out_len) {
  if ((!h_out && out_cap) || !out_len) return fail(BZG_ERR_INVALID, "null pointer");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  return copy_decoded(c, h_out, out_cap, out_len);
}

bzg_status bzg_decompress_device(const uint8_t* d_in, size_t len, uint8_t** d_out, size_t* out_len,
                                 void* stream) {
  if ((!d_in && len) || !d_out || !out_len) return fail(BZG_ERR_INVALID, "null pointer");
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  *d_out = nullptr;
  *out_len = 0;
  if ((st = decode_status(c, c->dec.run(d_in, len, s, &c->dlast))) != BZG_OK) return st;
  *d_out = c->dec.out();
  *out_len = c->dec.out_len();
  return BZG_OK;
}

bzg_status bzg_last_decode_times(float* h_ms5, int32_t* h_blocks, int32_t* h_streams) {
  bzg_status st;
  Ctx* c = ctx_for(g_dev, &st);
  if (!c) return st;
  std::lock_guard<std::mutex> lk(c->mu);
  const DecodeTimes& t = c->dlast;
  if (h_ms5) {
    h_ms5[0] = t.huff;
    h_ms5[1] = t.mtf;
    h_ms5[2] = t.sort;
    h_ms5[3] = t.walk;
    h_ms5[4] = t.rle1;
  }
  if (h_blocks) *h_blocks = t.blocks;
  if (h_streams) *h_streams = t.streams;
  return BZG_OK;
}

}  // extern "C"


// ------------------------------------------------------------ stage level
namespace {

// first HIP error of a sequence of enqueue calls
struct Ck {
  int e = 0;
  void operator()(hipError_t x) {
    if (!e && x != hipSuccess) e = (int)x;
  }
};
thread_local Ck ck;

struct StageLock {
  Ctx* c = nullptr;
  std::unique_lock<std::mutex> lk;
  hipStream_t s = nullptr;
};

bzg_status stage_begin(StageLock& L, void* stream, int32_t len) {
  ck.e = 0;
  bzg_status st;
  L.c = ctx_for(g_dev, &st);
  if (!L.c) return st;
  L.lk = std::unique_lock<std::mutex>(L.c->mu);
  if (hipSetDevice(L.c->device) != hipSuccess) return fail(BZG_ERR_HIP, "hipSetDevice failed");
  L.s = stream ? (hipStream_t)stream : L.c->stream;
  int e = L.c->stage.setup_single(len < 1 ? 1 : len, L.s);
  if (e) return hip_fail(e, "stage setup");
  return BZG_OK;
}

template <typename T>
int set_field(const Dev& d, size_t offset, T v, hipStream_t s) {
  return (int)hipMemcpyAsync(reinterpret_cast<char*>(d.st) + offset, &v, sizeof(T), hipMemcpyHostToDevice, s);
}

int sync(hipStream_t s) {
  int e = ck.e;
  ck.e = 0;
  int r = (int)hipStreamSynchronize(s);
  return e ? e : r;
}

}  // namespace

extern "C" {

int32_t bzg_rle1_bound(int32_t n) { return n < 0 ? 0 : cap_for(n); }

bzg_status bzg_crc32(const uint8_t* d_in, int32_t n, uint32_t* h_crc, void* stream) {
  if (!h_crc || n < 0 || (n > 0 && !d_in)) return fail(BZG_ERR_INVALID, "bad argument");
  if (n > kStageMax) return fail(BZG_ERR_INVALID, "n exceeds the stage limit (16 MiB)");
  if (n == 0) { *h_crc = 0; return BZG_OK; }
  StageLock L;
  bzg_status st = stage_begin(L, stream, n);
  if (st) return st;
  Engine& E = L.c->stage;
  std::vector<StreamSpec> ss{StreamSpec{0, (uint64_t)n, 9, n}};
  int e;
  if ((e = E.plan(ss, true)) || (e = E.setup(d_in, L.s))) return hip_fail(e, "stage setup");
  if ((e = E.run_front(L.s))) return hip_fail(e, "crc");
  uint32_t crc;
  ck(hipMemcpyAsync(&crc, &E.dev().st[0].crc, 4, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "crc");
  *h_crc = crc;
  return BZG_OK;
}

bzg_status bzg_rle1(const uint8_t* d_in, int32_t n, uint8_t* d_out, int32_t* h_out_len, void* stream) {
  if (!h_out_len || n < 0 || (n > 0 && (!d_in || !d_out))) return fail(BZG_ERR_INVALID, "bad argument");
  if (n > kStageMax) return fail(BZG_ERR_INVALID, "n exceeds the stage limit (16 MiB)");
  if (n == 0) { *h_out_len = 0; return BZG_OK; }  // rle1.cu:61-62
  StageLock L;
  bzg_status st = stage_begin(L, stream, n);
  if (st) return st;
  Engine& E = L.c->stage;
  std::vector<StreamSpec> ss{StreamSpec{0, (uint64_t)n, 9, n}};
  int e;
  if ((e = E.plan(ss, true)) || (e = E.setup(d_in, L.s))) return hip_fail(e, "stage setup");
  if ((e = E.run_front(L.s))) return hip_fail(e, "rle1");
  int32_t m;
  ck(hipMemcpyAsync(&m, &E.dev().st[0].m, 4, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "rle1");
  if (m > 0) ck(hipMemcpyAsync(d_out, E.dev().T, (size_t)m, hipMemcpyDeviceToDevice, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "rle1");
  *h_out_len = m;
  return BZG_OK;
}

bzg_status bzg_bwt(const uint8_t* d_in, int32_t m, int32_t* d_sa, int32_t* h_rounds, void* stream) {
  if (m < 0 || (m > 0 && (!d_in || !d_sa))) return fail(BZG_ERR_INVALID, "bad argument");
  if (m > kStageMax) return fail(BZG_ERR_INVALID, "m exceeds the stage limit (16 MiB)");
  if (m == 0) { if (h_rounds) *h_rounds = 0; return BZG_OK; }  // fbwt.cu:48-49
  StageLock L;
  bzg_status st = stage_begin(L, stream, m);
  if (st) return st;
  Engine& E = L.c->stage;
  const Dev& d = E.dev();
  int e;
  ck(hipMemcpyAsync(d.T, d_in, (size_t)m, hipMemcpyDeviceToDevice, L.s));
  if ((e = set_field<int32_t>(d, offsetof(BlockState, m), m, L.s))) return hip_fail(e, "bwt");
  int rounds = 0;
  if ((e = E.run_bwt(L.s, &rounds))) return hip_fail(e, "bwt");
  launch_sa_strip(d, d_sa, L.s);
  if ((e = sync(L.s))) return hip_fail(e, "bwt");
  if (h_rounds) *h_rounds = rounds;
  return BZG_OK;
}

bzg_status bzg_symbols_table(const uint8_t* d_in, int32_t n, uint8_t* h_table, int32_t* h_count, void* stream) {
  if (!h_table || !h_count || n < 0 || (n > 0 && !d_in)) return fail(BZG_ERR_INVALID, "bad argument");
  if (n > kStageMax) return fail(BZG_ERR_INVALID, "n exceeds the stage limit (16 MiB)");
  StageLock L;
  bzg_status st = stage_begin(L, stream, n);
  if (st) return st;
  Engine& E = L.c->stage;
  const Dev& d = E.dev();
  int e;
  if (n) ck(hipMemcpyAsync(d.T, d_in, (size_t)n, hipMemcpyDeviceToDevice, L.s));
  if ((e = set_field<int32_t>(d, offsetof(BlockState, m), n, L.s))) return hip_fail(e, "symbols");
  launch_presence(d, L.s);
  uint32_t pres[8];
  ck(hipMemcpyAsync(pres, d.st[0].present, 32, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "symbols");
  int cnt = 0;
  for (int i = 0; i < 256; i++)
    if ((pres[i >> 5] >> (i & 31)) & 1u) h_table[cnt++] = (uint8_t)i;
  *h_count = cnt;
  return BZG_OK;
}

bzg_status bzg_mtf(const uint8_t* d_rle1, const int32_t* d_sa, int32_t m, uint8_t* d_out, int32_t* h_orig_ptr,
                   uint8_t* h_present, int32_t* h_n_in_use, void* stream) {
  if (m < 0 || (m > 0 && (!d_rle1 || !d_sa || !d_out))) return fail(BZG_ERR_INVALID, "bad argument");
  if (m > kStageMax) return fail(BZG_ERR_INVALID, "m exceeds the stage limit (16 MiB)");
  if (m == 0) {  // fmtf.cu:150-151
    if (h_n_in_use) *h_n_in_use = 0;
    return BZG_OK;
  }
  StageLock L;
  bzg_status st = stage_begin(L, stream, m);
  if (st) return st;
  Engine& E = L.c->stage;
  const Dev& d = E.dev();
  int e;
  ck(hipMemcpyAsync(d.T, d_rle1, (size_t)m, hipMemcpyDeviceToDevice, L.s));
  if ((e = set_field<int32_t>(d, offsetof(BlockState, m), m, L.s))) return hip_fail(e, "mtf");
  launch_sa_attach(d, d_sa, L.s);
  launch_presence(d, L.s);
  launch_mtf(d, L.s);
  if ((e = (int)hipGetLastError())) return hip_fail(e, "mtf");
  ck(hipMemcpyAsync(d_out, d.MTF, (size_t)m, hipMemcpyDeviceToDevice, L.s));
  BlockState bs;
  ck(hipMemcpyAsync(&bs, d.st, sizeof bs, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "mtf");
  if (h_orig_ptr) *h_orig_ptr = bs.orig_ptr;
  if (h_n_in_use) *h_n_in_use = bs.n_in_use;
  if (h_present)
    for (int i = 0; i < 256; i++) h_present[i] = (bs.present[i >> 5] >> (i & 31)) & 1u;
  return BZG_OK;
}

bzg_status bzg_rle2(const uint8_t* d_mtf, int32_t m, int32_t table_size, uint16_t* d_out, uint32_t* d_out_len,
                    void* stream) {
  if (m < 0 || !d_out_len || (m > 0 && (!d_mtf || !d_out))) return fail(BZG_ERR_INVALID, "bad argument");
  if (m > kStageMax || table_size < 0 || table_size > 256) return fail(BZG_ERR_INVALID, "size out of range");
  StageLock L;
  if (m == 0) {  // rle2.cu:106-110
    bzg_status st;
    Ctx* c = ctx_for(g_dev, &st);
    if (!c) return st;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    ck(hipMemsetAsync(d_out_len, 0, 4, s));
    return sync(s) ? fail(BZG_ERR_HIP, "rle2") : BZG_OK;
  }
  bzg_status st = stage_begin(L, stream, m + 1);
  if (st) return st;
  Engine& E = L.c->stage;
  const Dev& d = E.dev();
  int e;
  ck(hipMemcpyAsync(d.MTF, d_mtf, (size_t)m, hipMemcpyDeviceToDevice, L.s));
  ck((hipError_t)set_field<int32_t>(d, offsetof(BlockState, m), m, L.s));
  if ((e = set_field<int32_t>(d, offsetof(BlockState, n_in_use), table_size, L.s))) return hip_fail(e, "rle2");
  launch_fnz(d, L.s);
  launch_rle2(d, L.s);
  if ((e = (int)hipGetLastError())) return hip_fail(e, "rle2");
  int32_t s_len;
  ck(hipMemcpyAsync(&s_len, &d.st[0].s, 4, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "rle2");
  ck(hipMemcpyAsync(d_out, d.SYM, (size_t)s_len * 2, hipMemcpyDeviceToDevice, L.s));
  ck(hipMemcpyAsync(d_out_len, &d.st[0].s, 4, hipMemcpyDeviceToDevice, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "rle2");
  return BZG_OK;
}

bzg_status bzg_histogram(const uint16_t* d_sym, int32_t s, int32_t* d_freq, void* stream) {
  if (s < 0 || !d_freq || (s > 0 && !d_sym)) return fail(BZG_ERR_INVALID, "bad argument");
  if (s > kStageMax) return fail(BZG_ERR_INVALID, "s exceeds the stage limit (16 MiB)");
  if (s == 0) return BZG_OK;
  StageLock L;
  bzg_status st = stage_begin(L, stream, s);
  if (st) return st;
  Engine& E = L.c->stage;
  const Dev& d = E.dev();
  int e;
  ck(hipMemcpyAsync(d.SYM, d_sym, (size_t)s * 2, hipMemcpyDeviceToDevice, L.s));
  ck((hipError_t)set_field<int32_t>(d, offsetof(BlockState, s), s, L.s));
  launch_sym_hist(d, L.s);
  std::vector<int32_t> h(258), add(258);
  ck(hipMemcpyAsync(add.data(), d.freq, 258 * 4, hipMemcpyDeviceToHost, L.s));
  ck(hipMemcpyAsync(h.data(), d_freq, 258 * 4, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "histogram");
  for (int i = 0; i < 258; i++) h[i] += add[i];
  ck(hipMemcpyAsync(d_freq, h.data(), 258 * 4, hipMemcpyHostToDevice, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "histogram");
  return BZG_OK;
}

bzg_status bzg_huff_build(const uint16_t* d_sym, int32_t s, int32_t alphabet_size, uint8_t* h_len, int32_t* h_code,
                          uint8_t* h_sel, int32_t* h_n_groups, int32_t* h_n_sel, void* stream) {
  if (s <= 0 || !d_sym || !h_len || !h_code || !h_sel || !h_n_groups || !h_n_sel)
    return fail(BZG_ERR_INVALID, "bad argument");
  if (s > kStageMax || alphabet_size < 0 || alphabet_size > 256) return fail(BZG_ERR_INVALID, "size out of range");
  StageLock L;
  bzg_status st = stage_begin(L, stream, s);
  if (st) return st;
  Engine& E = L.c->stage;
  const Dev& d = E.dev();
  int e;
  ck(hipMemcpyAsync(d.SYM, d_sym, (size_t)s * 2, hipMemcpyDeviceToDevice, L.s));
  ck((hipError_t)set_field<int32_t>(d, offsetof(BlockState, s), s, L.s));
  ck((hipError_t)set_field<int32_t>(d, offsetof(BlockState, n_in_use), alphabet_size, L.s));
  launch_sym_hist(d, L.s);
  launch_huffman(d, L.s);
  if ((e = (int)hipGetLastError())) return hip_fail(e, "huff_build");
  BlockState bs;
  std::vector<uint8_t> ln(6 * 260);
  std::vector<uint32_t> cd(6 * 260);
  ck(hipMemcpyAsync(&bs, d.st, sizeof bs, hipMemcpyDeviceToHost, L.s));
  ck(hipMemcpyAsync(ln.data(), d.len, ln.size(), hipMemcpyDeviceToHost, L.s));
  ck(hipMemcpyAsync(cd.data(), d.code, cd.size() * 4, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "huff_build");
  ck(hipMemcpyAsync(h_sel, d.sel, (size_t)bs.n_sel, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "huff_build");
  for (int g = 0; g < 6; g++)
    for (int v = 0; v < 258; v++) {
      h_len[g * 258 + v] = ln[g * 260 + v];
      h_code[g * 258 + v] = (int32_t)cd[g * 260 + v];
    }
  *h_n_groups = bs.n_groups;
  *h_n_sel = bs.n_sel;
  return BZG_OK;
}

bzg_status bzg_huff_encode(const uint16_t* d_sym, int32_t s, const uint8_t* h_len, const int32_t* h_code,
                           const uint8_t* d_sel, int32_t n_sel, int32_t n_groups, uint32_t* d_words, int64_t* h_bits,
                           void* stream) {
  if (s <= 0 || !d_sym || !h_len || !h_code || !d_sel || !d_words || !h_bits) return fail(BZG_ERR_INVALID, "bad argument");
  if (s > kStageMax || n_groups < 1 || n_groups > 6 || n_sel != (s + 49) / 50)
    return fail(BZG_ERR_INVALID, "size out of range");
  // the encoder's longest code is 17 bits (huffman.cu:321); the payload
  // capacity of a block is sized for that
  for (int g = 0; g < n_groups; g++)
    for (int v = 0; v < 258; v++)
      if (h_len[g * 258 + v] > 17) return fail(BZG_ERR_INVALID, "code length above 17");
  StageLock L;
  bzg_status st = stage_begin(L, stream, s);
  if (st) return st;
  Engine& E = L.c->stage;
  const Dev& d = E.dev();
  int e;
  std::vector<uint8_t> ln(6 * 260, 0);
  std::vector<uint32_t> cd(6 * 260, 0);
  for (int g = 0; g < 6; g++)
    for (int v = 0; v < 258; v++) {
      ln[g * 260 + v] = h_len[g * 258 + v];
      cd[g * 260 + v] = (uint32_t)h_code[g * 258 + v];
    }
  ck(hipMemcpyAsync(d.SYM, d_sym, (size_t)s * 2, hipMemcpyDeviceToDevice, L.s));
  ck(hipMemcpyAsync(d.sel, d_sel, (size_t)n_sel, hipMemcpyDeviceToDevice, L.s));
  ck(hipMemcpyAsync(d.len, ln.data(), ln.size(), hipMemcpyHostToDevice, L.s));
  ck(hipMemcpyAsync(d.code, cd.data(), cd.size() * 4, hipMemcpyHostToDevice, L.s));
  ck((hipError_t)set_field<int32_t>(d, offsetof(BlockState, s), s, L.s));
  ck((hipError_t)set_field<int32_t>(d, offsetof(BlockState, n_sel), n_sel, L.s));
  ck((hipError_t)set_field<int32_t>(d, offsetof(BlockState, n_groups), n_groups, L.s));
  launch_check_codes(d, n_groups, L.s);
  int32_t bad = 0;
  ck(hipMemcpyAsync(&bad, &d.st[0].err, 4, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "huff_encode");
  if (bad) return fail(BZG_ERR_INVALID, "symbol >= 258 or selector >= n_groups");
  launch_pack(d, L.s);
  if ((e = (int)hipGetLastError())) return hip_fail(e, "huff_encode");
  int64_t bits;
  ck(hipMemcpyAsync(&bits, &d.st[0].bits, 8, hipMemcpyDeviceToHost, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "huff_encode");
  ck(hipMemcpyAsync(d_words, d.pay, (size_t)((bits + 31) / 32) * 4, hipMemcpyDeviceToDevice, L.s));
  if ((e = sync(L.s))) return hip_fail(e, "huff_encode");
  *h_bits = bits;
  return BZG_OK;
}

}  // extern "C"
