// engine.h - host side of the batch pipeline: plans a batch of blocks,
// owns the device workspace and enqueues the stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <array>
#include <functional>
#include <vector>

#include "pipeline.h"

namespace bzg {

// One output stream: raw bytes [in_off, in_off+len) of the device input.
// lead >= 0 makes it a segment of a longer stream instead (blocks only, see
// StreamState), with crc_in the combined CRC of the blocks before it.
struct StreamSpec {
  uint64_t in_off;
  uint64_t len;
  int level;        // header digit 1..9
  int block_bytes;  // raw bytes per block (900000 = the reference's cut)
  int lead = -1;
  uint32_t crc_in = 0;
};

// Per-block result copied back to the host (the reference's BlockData,
// src/compression/compress.cu:13-24).
struct BlockResult {
  uint32_t crc;
  int32_t m, s, orig_ptr, n_in_use, n_groups, n_sel, rounds;
  uint8_t present[256];
  int64_t bits, hdr_bits;
  std::vector<uint32_t> hdr;    // header words (hdr_bits)
  std::vector<uint32_t> words;  // payload words (bits)
};

// Timing of the last run, from HIP events on the run's stream (ms).
struct StageTimes {
  float front = 0, bwt = 0, mtf = 0, huff = 0, pack = 0, assemble = 0;
  int rounds = 0;
  int64_t bwt_elems = 0;  // rows keyed and sorted, summed over rounds
  int64_t small_rows = 0, large_rows = 0;  // of which in small / large groups
  int64_t large_pass_rows = 0;             // large-group rows x radix passes
};

class Engine;

// Most prefix-doubling rounds a suffix sort runs.
constexpr int kMaxRounds = kMaxSortRounds;

// Plain path: the blocks of one stream, each through its own pipeline run
// on `s`, results copied back and serialised on the host (the reference's
// compress_block + BitWriter structure).  `d_in` is device memory.
int compress_plain_device(Engine& eng, const uint8_t* d_in, size_t len, int level, int32_t block_bytes,
                          std::vector<uint8_t>& out, hipStream_t s);
// One block alone through a pipeline run on `s`, its result copied back
// (the reference's compress_block, src/compression/compress.cu:60-182).
int compress_one_block(Engine& eng, const uint8_t* d_block, int32_t n, int level, BlockResult& r, hipStream_t s);
// Host serialisation of one stream from its blocks' results, in order
// (the reference's BitWriter pass, src/compression/compress.cu:284-375).
void serialize_stream(int level, const BlockResult* blocks, size_t n_blocks, std::vector<uint8_t>& out);

// Stage groups other than the suffix sort run as captured HIP graphs
// (process-wide switch; BZG_NO_GRAPHS=1 starts with graphs off).
void set_stage_graphs(bool on);
bool stage_graphs();

// per-kernel HIP-event probe (one target kernel name at a time); while a
// target is set, stage groups run eagerly instead of as graphs
void probe_set(const char* kernel);
bool probe_active();
int probe_read(float* total_ms, int* launches);

class Engine {
 public:
  explicit Engine(int device = 0);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // Host-side plan of a batch.  Returns 0 or a hipError_t value.
  // allow_big lets one block exceed 900000 raw bytes (stage entry points).
  int plan(const std::vector<StreamSpec>& streams, bool allow_big = false);
  // Allocates (grow-only) and uploads the plan; d_in holds the raw bytes.
  // With `out` the streams are assembled there (planned_out_bytes() bytes,
  // 4-byte aligned) instead of in the workspace's own output area.
  int setup(const uint8_t* d_in, hipStream_t s, uint8_t* out = nullptr);
  // Device bytes the planned streams may take (worst case).
  int64_t planned_out_bytes() const { return out_cap_ + 64; }
  // Runs every pipeline in caller-provided device memory instead of a
  // workspace of its own: setup then allocates nothing and fails with
  // hipErrorOutOfMemory when a plan needs more than `bytes`
  // (needed_workspace() tells how much the last setup wanted).
  void use_fixed_workspace(void* base, size_t bytes);
  size_t needed_workspace() const { return need_; }
  // Creates the per-round timing events of up to `rounds` suffix-sort
  // rounds now, so that runs with at most that many create none.
  int reserve_round_events(int rounds);
  // Full pipeline on the planned batch.  With assemble = true the .bz2
  // streams are written to the device output buffer.
  int run(hipStream_t s, bool assemble, StageTimes* times = nullptr);
  // Called with true before a run's front end is enqueued (a nonzero
  // return aborts the run) and with false once its suffix sort has ended on
  // the device; a caller that overlaps engines orders their sorts with it.
  void set_sort_gate(std::function<int(bool)> gate) { gate_ = std::move(gate); }
  // Whether a run of the planned batch would time autotuner variants.
  bool will_tune() const;
  int mtf_run_class() const;
  // Stage groups (used by run and by the per-stage entry points).
  int run_front(hipStream_t s);
  int run_bwt(hipStream_t s, int* rounds_out = nullptr);
  int run_mtf_rle2(hipStream_t s);
  int run_huffman(hipStream_t s);
  int run_pack(hipStream_t s);
  int run_assemble(hipStream_t s);

  // Single-block helper for the per-stage entry points: a one-block batch of
  // capacity for `len` elements (no raw input).
  int setup_single(int32_t len, hipStream_t s);

  // Copies per-stream layout (byte_off, bytes) and the batch totals of the
  // last run (totals()) to the host.
  int stream_layout(std::vector<StreamState>& out, hipStream_t s);
  const int64_t* totals() const { return totals_; }
  // Copies one block's results to the host.
  int block_result(int b, BlockResult& r, hipStream_t s);

  const Dev& dev() const { return d_; }
  int32_t nblk() const { return (int32_t)desc_.size(); }
  int32_t nstream() const { return (int32_t)ss_.size(); }
  uint8_t* out_buffer() const { return d_.out; }
  int64_t out_capacity() const { return d_.out_cap; }
  const std::vector<BlockDesc>& desc() const { return desc_; }
  // suffix-sort rows and GPU milliseconds of every round of the last run
  const std::vector<int64_t>& round_rows() const { return round_rows_; }
  const std::vector<float>& round_ms() const { return round_ms_; }
  int device() const { return device_; }
  // graph captures and replays of stage group i (0 front, 1 MTF+RLE2,
  // 2 Huffman, 3 pack, 4 assembly, 5 suffix-sort round 0, 6 later round)
  static constexpr int kStageGroups = 7;
  int64_t graph_captures(int i) const { return sgraph_[i].captures; }
  // per-engine switch (the plain per-block path plans a new batch per block,
  // where a capture per block would not pay off)
  void set_graphs(bool on) { use_graphs_ = on; }
  bool graphs() const { return use_graphs_; }
  int64_t graph_replays(int i) const { return sgraph_[i].replays; }

 private:
  int grow(size_t bytes, hipStream_t s);
  int run_stage(int idx, hipStream_t s, int (Engine::*fn)(hipStream_t), int variant = 0);
  // one suffix-sort round from device-side counts (round 0 / a later round)
  bool shape_tuning_pending() const;
  int sort_launch_class() const;
  int sort_round0(hipStream_t s);
  int sort_round(hipStream_t s);
  // Captured graphs of a stage group, one per recent device-side plan (a
  // batch split into sub-batches cycles through several plans per call).
  static constexpr int kGraphCache = 8;
  struct StageGraph {
    hipGraphExec_t exec[kGraphCache] = {};
    Dev key[kGraphCache] = {};
    int variant[kGraphCache] = {};   // launch choices captured with it
    int64_t used[kGraphCache] = {};  // last use, for eviction
    int64_t clock = 0;
    int64_t captures = 0, replays = 0;
  };
  StageGraph sgraph_[kStageGroups];
  // suffix sort: persistent grids, key bits and scatter shapes of the run
  SortGrids grids_{}, late_grids_{};
  bool grids_ready_ = false;
  int late_bits_ = 21;
  int shape0_ = 256, shape1_ = 256;
  bool late_large_seen_ = true;  // a later round had large groups (last timed run)
  bool use_graphs_ = true;
  int device_;
  Dev d_{};
  void* ws_ = nullptr;
  size_t ws_bytes_ = 0;
  bool fixed_ws_ = false;
  size_t need_ = 0;
  int64_t totals_[kTotals] = {};
  int32_t* h_ctr_ = nullptr;  // pinned
  std::vector<BlockDesc> desc_;
  std::vector<StreamState> ss_;
  std::vector<int32_t> tn_blk_, tc_blk_;
  int64_t total_cap_ = 0, sel_tot_ = 0, w_tot_ = 0, out_cap_ = 0;
  int32_t max_cap_ = 0;
  int64_t active_elems_ = 0;
  int64_t small_rows_ = 0, large_rows_ = 0, large_pass_rows_ = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> rev_;
  std::vector<std::array<hipEvent_t, 5>> tev_;  // around each radix pass, per round
  std::vector<int64_t> round_rows_;
  std::vector<float> round_ms_;
  hipEvent_t ev_[10] = {};  // stage bounds [0..6], BWT stage [7..8]
  int32_t h_ctr_prev_[CTR_N] = {};  // counters of a round whose radix passes are timed
  int init_err_ = 0;
  std::function<int(bool)> gate_;
};

}  // namespace bzg
