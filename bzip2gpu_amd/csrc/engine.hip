// engine.hip - batch planning, workspace and stage sequencing (host code).
#include "engine.h"
#include "tuner.h"

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>

namespace bzg {

namespace {
// BZG_TRACE=1: synchronise after every stage and print a timestamped line
// to stderr (diagnostics only; never set in timed runs).
bool trace_on() {
  static const bool on = [] {
    const char* v = getenv("BZG_TRACE");
    return v && v[0] == '1';
  }();
  return on;
}
void trace(hipStream_t s, const char* what, long a = -1, long b = -1, long c = -1) {
  if (!trace_on()) return;
  static auto t0 = std::chrono::steady_clock::now();
  hipError_t e = hipStreamSynchronize(s);
  double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  fprintf(stderr, "[bzg %10.3f ms] %s %ld %ld %ld (%s)\n", ms, what, a, b, c, hipGetErrorString(e));
  fflush(stderr);
}
}  // namespace

// ------------------------------------------------------------------ graphs
namespace {
std::atomic<int> g_graphs{-1};  // -1: from the environment (BZG_NO_GRAPHS=1 turns off)
}
void set_stage_graphs(bool on) { g_graphs = on ? 1 : 0; }
bool stage_graphs() {
  int v = g_graphs.load();
  if (v < 0) {
    const char* e = getenv("BZG_NO_GRAPHS");
    v = (e && e[0] == '1') ? 0 : 1;
    g_graphs = v;
  }
  return v == 1 && !trace_on();
}

// ------------------------------------------------------------------ probe
// Launches of the probed kernel are timed with an event pair each; host
// threads driving engines concurrently share the list (under a mutex) and
// each keeps the slot it has open.
namespace {
struct ProbeState {
  std::mutex mu;
  std::string target;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t used = 0;
};
ProbeState& probe() {
  static ProbeState p;
  return p;
}
thread_local long t_probe_slot = -1;  // this thread's open slot
}  // namespace

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

void probe_begin(const char* kernel, hipStream_t s) {
  ProbeState& P = probe();
  std::lock_guard<std::mutex> lk(P.mu);
  t_probe_slot = -1;
  if (P.target.empty() || P.target != kernel) return;
  if (capturing(s)) return;  // kernels inside captured stage graphs are not probed
  if (P.used == P.ev.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
    P.ev.emplace_back(a, b);
  }
  if (hipEventRecord(P.ev[P.used].first, s) != hipSuccess) return;
  // the slot is taken now; its second event follows the launch
  t_probe_slot = (long)P.used++;
  (void)hipEventRecord(P.ev[t_probe_slot].second, s);
}

void probe_end(const char* kernel, hipStream_t s) {
  ProbeState& P = probe();
  std::lock_guard<std::mutex> lk(P.mu);
  if (t_probe_slot < 0 || P.target != kernel) return;
  (void)hipEventRecord(P.ev[t_probe_slot].second, s);
  t_probe_slot = -1;
}

bool probe_active() {
  ProbeState& P = probe();
  std::lock_guard<std::mutex> lk(P.mu);
  return !P.target.empty();
}

// Whether the probed kernel may launch inside stage group `idx` (such a
// group runs eagerly so that its launches are timed; kernels of unknown
// stage make every group eager).
bool probe_in_stage(int idx) {
  std::string k;
  {
    ProbeState& P = probe();
    std::lock_guard<std::mutex> lk(P.mu);
    k = P.target;
  }
  if (k.empty()) return false;
  static const struct { const char* kernel; int stage; } kStage[] = {
      {"k_front1", 0}, {"k_rle1_write", 0}, {"k_mtf_gather", 1}, {"k_mtf_apply", 1}, {"k_rle2_write", 1},
      {"k_huff_build", 2}, {"k_huff_codes", 2}, {"k_pack", 3}, {"k_place", 4},
      {"k_lscatter", 5}, {"k_lscatter", 6}, {"k_keys", 6}, {"k_sort_small", 5}, {"k_sort_small", 6},
      {"k_lheads", 5}, {"k_lheads", 6}};
  bool known = false;
  for (const auto& e : kStage) {
    if (k != e.kernel) continue;
    known = true;
    if (e.stage == idx) return true;
  }
  return !known;
}

void probe_set(const char* kernel) {
  ProbeState& P = probe();
  std::lock_guard<std::mutex> lk(P.mu);
  P.target = kernel ? kernel : "";
  P.used = 0;
}

int probe_read(float* total_ms, int* launches) {
  ProbeState& P = probe();
  std::lock_guard<std::mutex> lk(P.mu);
  float tot = 0;
  for (size_t i = 0; i < P.used; i++) {
    hipError_t e = hipEventSynchronize(P.ev[i].second);
    if (e != hipSuccess) return (int)e;
    float ms = 0;
    e = hipEventElapsedTime(&ms, P.ev[i].first, P.ev[i].second);
    if (e != hipSuccess) return (int)e;
    tot += ms;
  }
  *total_ms = tot;
  *launches = (int)P.used;
  P.used = 0;
  return 0;
}

void launch_tile_scan_c(const Dev& d, hipStream_t s);

#define BZG_TRY(x)                       \
  do {                                   \
    hipError_t e_ = (x);                 \
    if (e_ != hipSuccess) return (int)e_; \
  } while (0)

Engine::Engine(int device) : device_(device) {
  int e = (int)hipSetDevice(device_);
  if (!e) e = (int)hipHostMalloc((void**)&h_ctr_, CTR_N * sizeof(int32_t), hipHostMallocDefault);
  for (auto& ev : ev_)
    if (!e) e = (int)hipEventCreate(&ev);
  init_err_ = e;
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (ws_ && !fixed_ws_) {  // allocated stream-ordered (grow)
    (void)hipFreeAsync(ws_, nullptr);
    (void)hipStreamSynchronize(nullptr);
  }
  if (h_ctr_) (void)hipHostFree(h_ctr_);
  for (auto& ev : ev_)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& pr : rev_) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  for (auto& ev : tev_)
    for (auto e : ev) (void)hipEventDestroy(e);
  for (auto& g : sgraph_)
    for (auto x : g.exec)
      if (x) (void)hipGraphExecDestroy(x);
}

int Engine::plan(const std::vector<StreamSpec>& streams, bool allow_big) {
  if (init_err_) return init_err_;
  desc_.clear();
  ss_.clear();
  tn_blk_.clear();
  tc_blk_.clear();
  total_cap_ = sel_tot_ = w_tot_ = out_cap_ = 0;
  max_cap_ = 0;
  int32_t tn = 0, tc = 0;
  for (size_t si = 0; si < streams.size(); si++) {
    const StreamSpec& S = streams[si];
    int level = std::min(9, std::max(1, S.level));
    int bb = S.block_bytes > 0 ? S.block_bytes : 900000;
    if (bb > (allow_big ? kStageMax : kMaxBlock)) return (int)hipErrorInvalidValue;
    StreamState st{};
    st.first_blk = (int32_t)desc_.size();
    st.level = level;
    if (S.lead > 7) return (int)hipErrorInvalidValue;
    st.lead = S.lead < 0 ? -1 : S.lead;
    st.crc_in = S.crc_in;
    int64_t sbytes = 16;
    for (uint64_t off = 0; off < S.len; off += (uint64_t)bb) {
      BlockDesc D{};
      D.in_off = S.in_off + off;
      D.n = (int32_t)std::min<uint64_t>((uint64_t)bb, S.len - off);
      D.cap = cap_for(D.n);
      D.off = (uint64_t)total_cap_;
      D.sel_off = (uint64_t)sel_tot_;
      D.w_off = (uint64_t)w_tot_;
      D.tile_n0 = tn;
      D.tile_c0 = tc;
      D.stream = (int32_t)si;
      D.level = level;
      int32_t b = (int32_t)desc_.size();
      for (int t = 0; t < tiles_for(D.n); t++) tn_blk_.push_back(b);
      for (int t = 0; t < tiles_for(D.cap); t++) tc_blk_.push_back(b);
      tn += tiles_for(D.n);
      tc += tiles_for(D.cap);
      total_cap_ += D.cap;
      sel_tot_ += sel_cap_for(D.cap);
      w_tot_ += pay_words_for(D.cap);
      max_cap_ = std::max(max_cap_, D.cap);
      sbytes += (pay_words_for(D.cap) + hdr_words_for(D.cap)) * 4;
      desc_.push_back(D);
    }
    st.n_blk = (int32_t)desc_.size() - st.first_blk;
    ss_.push_back(st);
    out_cap_ += (sbytes + 3) & ~3ll;
  }
  return 0;
}

// Stream-ordered growth: the old workspace is released and the new one
// allocated in order on `s` (every earlier use of the workspace was on a
// stream the caller synchronised before this call), so that one engine's
// growth does not wait for other engines' work on the device.
int Engine::grow(size_t bytes, hipStream_t s) {
  need_ = bytes;
  if (bytes <= ws_bytes_) return 0;
  if (fixed_ws_) return (int)hipErrorOutOfMemory;
  if (ws_) {
    BZG_TRY(hipFreeAsync(ws_, s));
    ws_ = nullptr;
    ws_bytes_ = 0;
  }
  size_t want = bytes + bytes / 8;
  BZG_TRY(hipMallocAsync(&ws_, want, s));
  ws_bytes_ = want;
  return 0;
}

void Engine::use_fixed_workspace(void* base, size_t bytes) {
  ws_ = base;
  ws_bytes_ = bytes;
  fixed_ws_ = true;
}

int Engine::reserve_round_events(int rounds) {
  while ((int)rev_.size() < rounds) {
    hipEvent_t a = nullptr, b = nullptr;
    BZG_TRY(hipEventCreate(&a));
    BZG_TRY(hipEventCreate(&b));
    rev_.emplace_back(a, b);
  }
  while ((int)tev_.size() < rounds) {
    std::array<hipEvent_t, 5> ev{};
    for (auto& e : ev) BZG_TRY(hipEventCreate(&e));
    tev_.push_back(ev);
  }
  return 0;
}

namespace {
struct Carver {
  size_t off = 0;
  char* base = nullptr;
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T) + 64;
    return p;
  }
};
}  // namespace


int Engine::setup(const uint8_t* d_in, hipStream_t s, uint8_t* out) {
  BZG_TRY(hipSetDevice(device_));
  const int64_t B = (int64_t)desc_.size();
  const int64_t S = (int64_t)ss_.size();
  const int64_t NTn = (int64_t)tn_blk_.size(), NTc = (int64_t)tc_blk_.size();
  const int64_t CAP = total_cap_;
  const int64_t seg_cap = CAP / 2 + B + 16;
  const int64_t pack_cap = CAP / kHalf + B + 16;
  const int64_t ltile_cap = CAP / 512 + B + 16;
  const int64_t hdr_stride = B ? hdr_words_for(max_cap_) : 8;
  auto carve = [&](Carver& c) {
    Dev d{};
    d.desc = c.take<BlockDesc>(B + 1);
    d.st = c.take<BlockState>(B + 1);
    d.ss = c.take<StreamState>(S + 1);
    d.tn_blk = c.take<int32_t>(NTn + 1);
    d.tc_blk = c.take<int32_t>(NTc + 1);
    d.tn_crc = c.take<uint32_t>(NTn + 1);
    d.tn_fb = c.take<int32_t>(NTn + 1);
    d.tn_lb = c.take<int32_t>(NTn + 1);
    d.tn_cnt = c.take<int32_t>(NTn + 1);
    d.tn_off = c.take<int32_t>(NTn + 1);
    d.tc_cnt = c.take<int32_t>(NTc + 1);
    d.tc_off = c.take<int32_t>(NTc + 1);
    d.tc_fnz = c.take<int32_t>(NTc + 1);
    d.lo = c.take<int32_t>((NTc + 1) * 256);
    d.T = c.take<uint8_t>(CAP + 64);
    d.SA = c.take<uint32_t>(CAP + 16);
    d.RK = c.take<uint32_t>(CAP + 16);
    d.RKB = c.take<uint64_t>(CAP + 16);
    d.rkb_cnt = c.take<int32_t>(NTc + 1);
    d.KEY = c.take<uint32_t>(CAP + 16);
    d.SA2 = c.take<uint32_t>(CAP + 16);
    d.KEY2 = c.take<uint32_t>(CAP + 16);
    d.MTF = c.take<uint8_t>(CAP + 64);
    d.SYM = c.take<uint16_t>(CAP + 64);
    d.freq = c.take<int32_t>((B + 1) * 260);
    d.len = c.take<uint8_t>((B + 1) * 6 * 260);
    d.code = c.take<uint32_t>((B + 1) * 6 * 260);
    d.sel = c.take<uint8_t>(sel_tot_ + 64);
    d.selmtf = c.take<uint8_t>(sel_tot_ + 64);
    d.gbits = c.take<int32_t>(sel_tot_ + 64);
    d.pay = c.take<uint32_t>(w_tot_ + 64);
    d.hdr = c.take<uint32_t>((B + 1) * hdr_stride);
    d.seg[0] = c.take<Seg>(seg_cap);
    d.seg[1] = c.take<Seg>(seg_cap);
    d.pa = c.take<int32_t>(seg_cap);
    d.pb = c.take<int32_t>(seg_cap);
    d.pc = c.take<int32_t>(seg_cap);
    d.packs = c.take<int32_t>(pack_cap + 2);
    d.horder = c.take<int32_t>(B + 2);
    d.rk_list = c.take<int32_t>(B + 2);
    d.rpk = c.take<int32_t>(pack_cap + 2);
    d.ltile_seg = c.take<int32_t>(ltile_cap);
    d.ltile_pos = c.take<int32_t>(ltile_cap);
    d.llist = c.take<int32_t>(ltile_cap);
    d.hist = c.take<uint32_t>(ltile_cap * 256);
    d.lsum = c.take<int32_t>(ltile_cap * 2);
    d.ctr = c.take<int32_t>(CTR_N);
    d.round_rows = c.take<int64_t>(2 * kMaxSortRounds);
    d.totals = c.take<int64_t>(kTotals);
    d.scan_part = c.take<int32_t>(3 * kScanWG);
    if (!out) d.out = c.take<uint8_t>(out_cap_ + 64);
    return d;
  };
  Carver probe;
  carve(probe);
  BZG_TRY((hipError_t)grow(probe.off + 4096, s));
  Carver real;
  real.base = static_cast<char*>(ws_);
  Dev d = carve(real);
  if (out) d.out = out;
  d.in = d_in;
  d.nblk = (int32_t)B;
  d.nstream = (int32_t)S;
  d.ntile_n = (int32_t)NTn;
  d.ntile_c = (int32_t)NTc;
  d.mtf_run = 4;
  d.tiles_max = tiles_for(max_cap_);
  d.total_cap = CAP;
  d.hdr_stride = hdr_stride;
  d.seg_cap = seg_cap;
  d.ltile_cap = ltile_cap;
  d.pack_cap = pack_cap;
  d.out_cap = out_cap_;
  d_ = d;
  if (B) {
    BZG_TRY(hipMemcpyAsync((void*)d.desc, desc_.data(), B * sizeof(BlockDesc), hipMemcpyHostToDevice, s));
    BZG_TRY(hipMemcpyAsync((void*)d.tn_blk, tn_blk_.data(), NTn * sizeof(int32_t), hipMemcpyHostToDevice, s));
    BZG_TRY(hipMemcpyAsync((void*)d.tc_blk, tc_blk_.data(), NTc * sizeof(int32_t), hipMemcpyHostToDevice, s));
  }
  if (S) BZG_TRY(hipMemcpyAsync(d.ss, ss_.data(), S * sizeof(StreamState), hipMemcpyHostToDevice, s));
  return 0;
}

int Engine::setup_single(int32_t len, hipStream_t s) {
  std::vector<StreamSpec> one{StreamSpec{0, (uint64_t)len, 9, len}};
  int rc = plan(one, true);
  if (rc) return rc;
  if (desc_.size() != 1) return (int)hipErrorInvalidValue;
  rc = setup(nullptr, s);
  if (rc) return rc;
  launch_block_init(d_, s);
  return (int)hipGetLastError();
}

int Engine::run_front(hipStream_t s) {
  launch_block_init(d_, s);
  launch_crc_rle1(d_, s);
  return (int)hipGetLastError();
}

// The radix-scatter shape is still being timed for round 0, or for the
// later rounds when the last timed run had large groups there.
bool Engine::shape_tuning_pending() const {
  Tuner& tu = Tuner::get();
  return tu.exploring(TUNE_RADIX_SHAPE, TUNE_CLASS_ROUND0) ||
         (late_large_seen_ && tu.exploring(TUNE_RADIX_SHAPE, TUNE_CLASS_LATER));
}

// Size class of the sort-launch knob: launch latency weighs on small
// batches, exact grids on large ones.
int Engine::sort_launch_class() const {
  return total_cap_ < (8 << 20) ? 0 : total_cap_ < (128 << 20) ? 1 : 2;
}

// Size class of the MTF-run knob: small batches need every segment's wave
// to fill the device, large ones can take longer runs.
int Engine::mtf_run_class() const { return total_cap_ < (32 << 20) ? 0 : 1; }

bool Engine::will_tune() const {
  Tuner& tu = Tuner::get();
  return shape_tuning_pending() || tu.exploring(TUNE_SORT_LAUNCH, sort_launch_class()) ||
         tu.exploring(TUNE_MTF_RUN, mtf_run_class());
}

int Engine::sort_round0(hipStream_t s) {
  launch_bwt_round(d_, grids_, true, late_bits_, shape0_, s, nullptr, nullptr);
  return (int)hipGetLastError();
}

int Engine::sort_round(hipStream_t s) {
  launch_bwt_round(d_, late_grids_, false, late_bits_, shape1_, s, nullptr, nullptr);
  return (int)hipGetLastError();
}

// The suffix sort, in one of three launch policies (an autotuner knob, all
// bit-identical):
//  * per round: the round's plan runs, its totals are read back, and every
//    kernel is launched with one workgroup per work item (skipped when it
//    has none); the readback also ends the loop;
//  * device-driven: every round plans itself; round 0's kernels get grids
//    bounded by the plan and the later rounds' kernels persistent grids (one
//    or four residencies of the device) that take their work from device
//    memory, so round 0 and then chunks of later rounds are enqueued (or
//    replayed as captured graphs) back to back, with one readback per chunk
//    to see whether the group list has run empty; rounds past the last one
//    do nothing.
// While the tuner still times a radix-scatter shape, rounds run per round
// with events around the radix passes (the tuner's unit: rows x passes).
int Engine::run_bwt(hipStream_t s, int* rounds_out) {
  active_elems_ = 0;
  small_rows_ = large_rows_ = large_pass_rows_ = 0;
  round_rows_.clear();
  round_ms_.clear();
  if (!grids_ready_) {
    grids_ = sort_grids(device_);
    grids_ready_ = true;
    if (trace_on())
      fprintf(stderr, "[bzg] sort grids: light %d keys %d small %d/%d scatter %d/%d/%d heads %d apply %d\n",
              grids_.light, grids_.keys, grids_.small, grids_.small64, grids_.scatter[0], grids_.scatter[1],
              grids_.scatter[2], grids_.heads, grids_.apply);
  }
  if (const int rc = reserve_round_events(kMaxRounds)) return rc;
  late_bits_ = 1;
  while ((1 << late_bits_) < max_cap_ + 1) late_bits_++;
  const int late_pass = (late_bits_ + 7) / 8;
  Tuner& tu = Tuner::get();
  const int v0 = tu.pick(TUNE_RADIX_SHAPE, TUNE_CLASS_ROUND0), v1 = tu.pick(TUNE_RADIX_SHAPE, TUNE_CLASS_LATER);
  shape0_ = tu.value(TUNE_RADIX_SHAPE, v0);
  shape1_ = tu.value(TUNE_RADIX_SHAPE, v1);
  const bool shape_tuning = shape_tuning_pending();
  const int mode_cls = sort_launch_class();
  const int mode = shape_tuning ? 0 : tu.pick(TUNE_SORT_LAUNCH, mode_cls);
  // device-driven later rounds: persistent grids of 1 or 4 residencies
  late_grids_ = grids_;
  if (mode == 2) {
    SortGrids& g = late_grids_;
    for (int* v : {&g.light, &g.keys, &g.small, &g.small64, &g.scatter[0], &g.scatter[1], &g.scatter[2], &g.heads,
                   &g.apply})
      *v *= 4;
  }
  BZG_TRY(hipEventRecord(ev_[7], s));
  launch_bwt_init(d_, s);
  BZG_TRY(hipGetLastError());
  int launched = 0;
  if (mode == 0) {
    bool late_large = false;
    int timed = -1, timed_pass = 0;  // a round whose radix passes await their report
    for (;; launched++) {
      launch_bwt_plan(d_, s);
      BZG_TRY(hipMemcpyAsync(h_ctr_, d_.ctr, CTR_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      BZG_TRY(hipStreamSynchronize(s));
      if (timed >= 0) {  // the previous round is complete
        float ms = 0.f;
        const int64_t units = (int64_t)h_ctr_prev_[CTR_LELEMS] * timed_pass;
        if (units > 0 && hipEventElapsedTime(&ms, tev_[timed][0], tev_[timed][timed_pass]) == hipSuccess)
          tu.report(TUNE_RADIX_SHAPE, timed == 0 ? TUNE_CLASS_ROUND0 : TUNE_CLASS_LATER, timed == 0 ? v0 : v1,
                    (double)ms * 1e6, (double)units);
        timed = -1;
      }
      if (h_ctr_[CTR_ERR]) return (int)hipErrorIllegalState;
      if (h_ctr_[CTR_NSEG] == 0) break;
      if (launched >= kMaxRounds - 1) return (int)hipErrorIllegalState;
      const bool first = launched == 0;
      if (!first && h_ctr_[CTR_LTILES] > 0) late_large = true;
      trace(s, "bwt round start", launched, h_ctr_[CTR_NSEG], h_ctr_[CTR_LTILES]);
      const hipEvent_t* pass_ev = nullptr;
      if (shape_tuning && h_ctr_[CTR_LTILES] > 0) {
        pass_ev = tev_[launched].data();
        timed = launched;
        timed_pass = first ? 4 : late_pass;
        std::memcpy(h_ctr_prev_, h_ctr_, sizeof h_ctr_prev_);
      }
      BZG_TRY(hipEventRecord(rev_[launched].first, s));
      launch_bwt_round(d_, grids_, first, late_bits_, first ? shape0_ : shape1_, s, pass_ev, h_ctr_);
      BZG_TRY(hipGetLastError());
      BZG_TRY(hipEventRecord(rev_[launched].second, s));
    }
    if (shape_tuning) late_large_seen_ = late_large;
  } else {
    BZG_TRY(hipEventRecord(rev_[0].first, s));
    int rc = run_stage(5, s, &Engine::sort_round0, shape0_);
    if (rc) return rc;
    BZG_TRY(hipEventRecord(rev_[0].second, s));
    launched = 1;
    for (int chunk = 2;; chunk = 4) {
      BZG_TRY(hipMemcpyAsync(h_ctr_, d_.ctr, CTR_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      BZG_TRY(hipStreamSynchronize(s));
      if (h_ctr_[CTR_ERR]) return (int)hipErrorIllegalState;
      if (h_ctr_[launched & 1] == 0) break;  // the list of the next round is empty
      if (launched >= kMaxRounds) return (int)hipErrorIllegalState;
      for (int k = 0; k < chunk && launched < kMaxRounds; k++, launched++) {
        BZG_TRY(hipEventRecord(rev_[launched].first, s));
        rc = run_stage(6, s, &Engine::sort_round, shape1_ | mode << 16);
        if (rc) return rc;
        BZG_TRY(hipEventRecord(rev_[launched].second, s));
      }
      trace(s, "bwt rounds", launched);
    }
  }
  // rows of every round (small groups, then large groups) and round count
  int64_t rr[2 * kMaxRounds];
  BZG_TRY(hipEventRecord(ev_[8], s));
  BZG_TRY(hipMemcpyAsync(rr, d_.round_rows, sizeof rr, hipMemcpyDeviceToHost, s));
  BZG_TRY(hipMemcpyAsync(h_ctr_, d_.ctr, CTR_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  BZG_TRY(hipStreamSynchronize(s));
  if (h_ctr_[CTR_ERR]) return (int)hipErrorIllegalState;
  const int ran = std::min(h_ctr_[CTR_RAN], launched);
  for (int r = 0; r < ran; r++) {
    const int64_t sm = rr[r], lg = rr[kMaxRounds + r];
    small_rows_ += sm;
    large_rows_ += lg;
    // radix passes that read KEY / SA (round 0's first pass reads the text)
    large_pass_rows_ += lg * (r == 0 ? 3 : late_pass);
    round_rows_.push_back(sm + lg);
  }
  active_elems_ = small_rows_ + large_rows_;
  if (rounds_out) *rounds_out = ran;
  round_ms_.assign((size_t)ran, 0.f);
  for (int r = 0; r < ran; r++) BZG_TRY(hipEventElapsedTime(&round_ms_[r], rev_[r].first, rev_[r].second));
  // the launch policy's cost per sorted row (BWT stage, both policies)
  if (!shape_tuning && tu.state(TUNE_SORT_LAUNCH, mode_cls, nullptr, nullptr) < 0 && active_elems_ > 0) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev_[7], ev_[8]) == hipSuccess)
      tu.report(TUNE_SORT_LAUNCH, mode_cls, mode, (double)ms * 1e6, (double)active_elems_);
  }
  return 0;
}

int Engine::run_mtf_rle2(hipStream_t s) {
  launch_mtf_rle2(d_, s);
  return (int)hipGetLastError();
}

int Engine::run_huffman(hipStream_t s) {
  launch_huffman(d_, s);
  return (int)hipGetLastError();
}

int Engine::run_pack(hipStream_t s) {
  launch_pack(d_, s);
  return (int)hipGetLastError();
}

int Engine::run_assemble(hipStream_t s) {
  launch_assemble(d_, s);
  return (int)hipGetLastError();
}

// Runs a stage group as a captured HIP graph.  A graph is instantiated once
// per device-side plan (the Dev struct: workspace, input and output
// pointers, block/tile/stream counts, hence every grid size) and replayed
// while that plan recurs; the stage groups have no host-sized launches,
// copies or memsets, so a replay does exactly what the eager launches do.
int Engine::run_stage(int idx, hipStream_t s, int (Engine::*fn)(hipStream_t), int variant) {
  // eager on the legacy stream, inside a caller's own capture, when off, or
  // while a kernel is probed (events cannot bracket a kernel inside a graph)
  if (!use_graphs_ || !stage_graphs() || s == nullptr || capturing(s) || probe_in_stage(idx)) return (this->*fn)(s);
  StageGraph& g = sgraph_[idx];
  g.clock++;
  int slot = -1, victim = 0;
  for (int i = 0; i < kGraphCache; i++) {
    if (g.exec[i] && g.variant[i] == variant && std::memcmp(&g.key[i], &d_, sizeof(Dev)) == 0) {
      slot = i;
      break;
    }
    if (!g.exec[i] || (g.exec[victim] && g.used[i] < g.used[victim])) victim = i;
  }
  if (slot < 0) {
    slot = victim;
    if (g.exec[slot]) (void)hipGraphExecDestroy(g.exec[slot]);
    g.exec[slot] = nullptr;
    hipGraph_t graph = nullptr;
    BZG_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = (this->*fn)(s);
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc ? rc : (int)e;
    }
    const hipError_t ei = hipGraphInstantiate(&g.exec[slot], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      g.exec[slot] = nullptr;
      return (int)ei;
    }
    std::memcpy(&g.key[slot], &d_, sizeof(Dev));
    g.variant[slot] = variant;
    g.captures++;
  }
  g.used[slot] = g.clock;
  g.replays++;
  return (int)hipGraphLaunch(g.exec[slot], s);
}

int Engine::run(hipStream_t s, bool assemble, StageTimes* times) {
  BZG_TRY(hipSetDevice(device_));
  if (desc_.empty()) {
    if (assemble && !ss_.empty()) {
      int rc = run_assemble(s);
      trace(s, "assemble (no blocks) done");
      return rc;
    }
    return 0;
  }
  int rounds = 0;
  trace(s, "run start", (long)desc_.size(), (long)ss_.size());
  int rc = gate_ ? gate_(true) : 0;
  if (rc) return rc;
  // front end and suffix sort (the gate is left once the sort has ended:
  // run_bwt returns after its last readback)
  rc = times ? (int)hipEventRecord(ev_[0], s) : 0;
  if (!rc) rc = run_stage(0, s, &Engine::run_front);
  trace(s, "front done");
  if (!rc && times) rc = (int)hipEventRecord(ev_[1], s);
  if (!rc) rc = run_bwt(s, &rounds);
  if (gate_) (void)gate_(false);
  if (rc) return rc;
  trace(s, "bwt done", rounds);
  // MTF segments per wave: an autotuner knob, timed over the MTF + RLE2
  // stage (unit: capacity tiles) while it is explored
  Tuner& tu = Tuner::get();
  const int mcls = mtf_run_class();
  const int mv = tu.pick(TUNE_MTF_RUN, mcls);
  const bool mtf_timing = tu.exploring(TUNE_MTF_RUN, mcls);
  d_.mtf_run = tu.value(TUNE_MTF_RUN, mv);
  if (times || mtf_timing) BZG_TRY(hipEventRecord(ev_[2], s));
  rc = run_stage(1, s, &Engine::run_mtf_rle2);
  if (rc) return rc;
  trace(s, "mtf+rle2 done");
  if (times || mtf_timing) BZG_TRY(hipEventRecord(ev_[3], s));
  rc = run_stage(2, s, &Engine::run_huffman);
  if (rc) return rc;
  trace(s, "huffman done");
  if (times) BZG_TRY(hipEventRecord(ev_[4], s));
  rc = run_stage(3, s, &Engine::run_pack);
  if (rc) return rc;
  trace(s, "pack done");
  if (times) BZG_TRY(hipEventRecord(ev_[5], s));
  if (assemble) {
    rc = run_stage(4, s, &Engine::run_assemble);
    if (rc) return rc;
    trace(s, "assemble done");
  }
  launch_block_totals(d_, s);
  if (times) {
    BZG_TRY(hipEventRecord(ev_[6], s));
    BZG_TRY(hipEventSynchronize(ev_[6]));
    float* f[6] = {&times->front, &times->bwt, &times->mtf, &times->huff, &times->pack, &times->assemble};
    for (int i = 0; i < 6; i++) BZG_TRY(hipEventElapsedTime(f[i], ev_[i], ev_[i + 1]));
    times->rounds = rounds;
    times->bwt_elems = active_elems_;
    times->small_rows = small_rows_;
    times->large_rows = large_rows_;
    times->large_pass_rows = large_pass_rows_;
  }
  if (mtf_timing) {
    float ms = 0.f;
    BZG_TRY(hipEventSynchronize(ev_[3]));
    if (hipEventElapsedTime(&ms, ev_[2], ev_[3]) == hipSuccess)
      tu.report(TUNE_MTF_RUN, mcls, mv, (double)ms * 1e6, (double)d_.ntile_c);
  }
  return 0;
}

int Engine::stream_layout(std::vector<StreamState>& out, hipStream_t s) {
  out.resize(ss_.size());
  for (auto& t : totals_) t = 0;
  if (ss_.empty()) return 0;
  BZG_TRY(hipMemcpyAsync(out.data(), d_.ss, ss_.size() * sizeof(StreamState), hipMemcpyDeviceToHost, s));
  if (!desc_.empty()) BZG_TRY(hipMemcpyAsync(totals_, d_.totals, sizeof totals_, hipMemcpyDeviceToHost, s));
  BZG_TRY(hipStreamSynchronize(s));
  return 0;
}

int Engine::block_result(int b, BlockResult& r, hipStream_t s) {
  BlockState st;
  BZG_TRY(hipMemcpyAsync(&st, d_.st + b, sizeof st, hipMemcpyDeviceToHost, s));
  BZG_TRY(hipStreamSynchronize(s));
  r.crc = st.crc; r.m = st.m; r.s = st.s; r.orig_ptr = st.orig_ptr;
  r.n_in_use = st.n_in_use; r.n_groups = st.n_groups; r.n_sel = st.n_sel;
  r.rounds = st.rounds; r.bits = st.bits; r.hdr_bits = st.hdr_bits;
  for (int i = 0; i < 256; i++) r.present[i] = (st.present[i >> 5] >> (i & 31)) & 1u;
  r.hdr.resize((size_t)((st.hdr_bits + 31) / 32));
  r.words.resize((size_t)((st.bits + 31) / 32));
  if (!r.hdr.empty())
    BZG_TRY(hipMemcpyAsync(r.hdr.data(), d_.hdr + (int64_t)b * d_.hdr_stride, r.hdr.size() * 4, hipMemcpyDeviceToHost, s));
  if (!r.words.empty())
    BZG_TRY(hipMemcpyAsync(r.words.data(), d_.pay + desc_[b].w_off, r.words.size() * 4, hipMemcpyDeviceToHost, s));
  BZG_TRY(hipStreamSynchronize(s));
  return 0;
}

namespace {
// MSB-first host bit writer (the reference's BitWriter, compress.cu:26-58).
struct HostBits {
  std::vector<uint8_t>& out;
  uint64_t buf = 0;
  int n = 0;
  explicit HostBits(std::vector<uint8_t>& o) : out(o) {}
  void put(uint32_t v, int k) {
    if (k == 0) return;
    v &= (uint32_t)((1ull << k) - 1);
    buf = (buf << k) | v;
    n += k;
    while (n >= 8) {
      n -= 8;
      out.push_back((uint8_t)(buf >> n));
    }
  }
  void put_words(const std::vector<uint32_t>& w, int64_t bits) {
    int64_t rem = bits;
    for (size_t i = 0; i < w.size() && rem > 0; i++) {
      int take = rem < 32 ? (int)rem : 32;
      put(w[i] >> (32 - take), take);
      rem -= take;
    }
  }
  void pad() {
    if (n > 0) put(0, 8 - n);
  }
};
}  // namespace

void serialize_stream(int level, const BlockResult* blocks, size_t n_blocks, std::vector<uint8_t>& out) {
  level = std::min(9, std::max(1, level));
  HostBits bw(out);
  bw.put(0x42, 8); bw.put(0x5A, 8); bw.put(0x68, 8); bw.put((uint32_t)(0x30 + level), 8);
  uint32_t combined = 0;
  for (size_t i = 0; i < n_blocks; i++) {
    const BlockResult& r = blocks[i];
    combined = ((combined << 1) | (combined >> 31)) ^ r.crc;
    bw.put_words(r.hdr, r.hdr_bits);
    bw.put_words(r.words, r.bits);
  }
  bw.put(0x17, 8); bw.put(0x72, 8); bw.put(0x45, 8); bw.put(0x38, 8); bw.put(0x50, 8); bw.put(0x90, 8);
  bw.put(combined, 32);
  bw.pad();
}

int compress_one_block(Engine& eng, const uint8_t* d_block, int32_t n, int level, BlockResult& r, hipStream_t s) {
  int e;
  std::vector<StreamSpec> ss{StreamSpec{0, (uint64_t)n, level, n}};
  if ((e = eng.plan(ss))) return e;
  if ((e = eng.setup(d_block, s))) return e;
  if ((e = eng.run(s, false, nullptr))) return e;
  return eng.block_result(0, r, s);
}

int compress_plain_device(Engine& eng, const uint8_t* d_in, size_t len, int level, int32_t block_bytes,
                          std::vector<uint8_t>& out, hipStream_t s) {
  level = std::min(9, std::max(1, level));
  out.clear();
  out.reserve(len / 2 + 64);
  // every block is a new one-block plan: the stage groups run eagerly
  struct GraphsOff {
    Engine& g;
    bool was;
    explicit GraphsOff(Engine& x) : g(x), was(x.graphs()) { g.set_graphs(false); }
    ~GraphsOff() { g.set_graphs(was); }
  } graphs_off(eng);
  std::vector<BlockResult> res((len + (size_t)block_bytes - 1) / (size_t)block_bytes);
  for (size_t k = 0; k < res.size(); k++) {
    const size_t off = k * (size_t)block_bytes;
    const size_t n = len - off < (size_t)block_bytes ? len - off : (size_t)block_bytes;
    const int e = compress_one_block(eng, d_in + off, (int32_t)n, level, res[k], s);
    if (e) return e;
  }
  serialize_stream(level, res.data(), res.size(), out);
  return 0;
}

}  // namespace bzg
