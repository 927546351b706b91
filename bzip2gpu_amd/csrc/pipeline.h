// pipeline.h - device views of one batch and the stage launchers.
//
// A batch is a list of blocks (each <= 900000 raw bytes) from one or more
// output streams.  Every stage kernel works on the whole batch at once: a
// batch of one block is the plain per-block path, a batch of many blocks is
// the batched path.  All buffers live in one device workspace; per-block
// arrays are indexed by BlockDesc::off (capacity-sized slots) so no stage
// needs a host round trip to place its output.
#pragma once

#include "common.h"

namespace bzg {

struct Seg {           // an unresolved group of the suffix sort
  int32_t blk;         // block index in the batch
  int32_t start;       // first SA index (block-local)
  int32_t len;         // number of rotations in the group (>= 2)
  int32_t pad;
};

constexpr int kPack = 2048;          // elements per small-group sort workgroup
constexpr int kHalf = kPack / 2;     // groups longer than this use the tile path
constexpr int kScanWG = 512;         // fixed grid of the seg-list scans
constexpr int kHdrFixedBits = 48 + 32 + 1 + 24 + 16 + 16 * 16 + 3 + 15;

__host__ __device__ inline int32_t sel_cap_for(int32_t cap) { return cap / kGSize + 2; }
__host__ __device__ inline int64_t pay_words_for(int32_t cap) { return (int64_t)cap * 17 / 32 + 8; }
__host__ __device__ inline int64_t hdr_words_for(int32_t cap) {
  return ((int64_t)sel_cap_for(cap) * 7 + kMaxGroups * (5 + kMaxAlpha * 33) + kHdrFixedBits) / 32 + 8;
}

struct Dev {
  // batch description
  const BlockDesc* desc;
  BlockState* st;
  StreamState* ss;
  const uint8_t* in;
  int32_t nblk, nstream;
  int32_t ntile_n, ntile_c;
  int32_t mtf_run;   // MTF segments per wave (1, 2, 4 or 8)
  int32_t tiles_max; // capacity tiles of the largest block
  int64_t total_cap;
  // tile -> block maps
  const int32_t* tn_blk;
  const int32_t* tc_blk;
  // raw-tile scratch
  uint32_t* tn_crc;
  int32_t* tn_fb;
  int32_t* tn_lb;
  int32_t* tn_cnt;
  int32_t* tn_off;
  // capacity-tile scratch
  int32_t* tc_cnt;
  int32_t* tc_off;
  int32_t* tc_fnz;
  int32_t* lo;         // [ntile_c][256] MTF last-occurrence summaries
  // capacity-sized block arrays
  uint8_t* T;          // RLE1 output
  uint32_t* SA;
  uint32_t* RK;
  uint64_t* RKB;       // bucketed (head << 32 | start) rank updates, kRkBucket starts per bucket
  int32_t* rkb_cnt;    // entries per bucket, indexed by the bucket's first capacity tile
  uint32_t* KEY;
  uint32_t* SA2;
  uint32_t* KEY2;
  uint8_t* MTF;
  uint16_t* SYM;
  // Huffman
  int32_t* freq;       // [nblk][260]
  uint8_t* len;        // [nblk][6][260]
  uint32_t* code;      // [nblk][6][260]
  uint8_t* sel;        // sel_off-indexed
  uint8_t* selmtf;
  int32_t* gbits;      // per capacity tile of a block: payload bits, then first bit (sel_off-indexed)
  uint32_t* pay;       // payload words (w_off-indexed)
  uint32_t* hdr;       // header words (h_off = per-block stride hdr_stride)
  int64_t hdr_stride;  // words per block in hdr
  // suffix sort work lists
  Seg* seg[2];
  int64_t seg_cap;
  int32_t* pa;         // per-seg exclusive prefix of small-group lengths
  int32_t* pb;         // per-seg exclusive prefix of large-group tile counts
  int32_t* pc;         // per-seg exclusive prefix of large-group count
  int32_t* packs;      // pack -> first seg index
  int32_t* horder;     // blocks in code-length build order (largest alphabet first)
  int32_t* rk_list;    // blocks whose suffix-sort ranks are built this round
  int32_t* rpk;        // round 1: the packs sorted by composite keys
  int32_t* ltile_seg;  // large tile -> seg index
  int32_t* ltile_pos;  // large tile -> first element within its seg
  int32_t* llist;      // compact list of large seg indices
  uint32_t* hist;      // [ltile][256] digit histograms / scatter offsets
  int32_t* lsum;       // [ltile][2] first/last group boundary in tile
  int64_t ltile_cap;
  int64_t pack_cap;
  int32_t* ctr;        // counters: [0],[1] list sizes; [2..] plan totals, round
  int64_t* round_rows; // [kMaxSortRounds] rows sorted in each suffix-sort round
  int64_t* totals;     // [kTotals] batch totals of the last run (k_block_totals)
  int32_t* scan_part;  // [3][kScanWG] scan partials
  // output
  uint8_t* out;
  int64_t out_cap;
};

// ---- per-kernel timing probe (HIP events around the launches of one named
// kernel on the launching stream; off unless a target name is set)
void probe_begin(const char* kernel, hipStream_t s);
void probe_end(const char* kernel, hipStream_t s);

// counter slots
enum : int {
  CTR_LIST0 = 0,
  CTR_LIST1 = 1,
  CTR_SMALL = 2,   // total elements in small groups
  CTR_LTILES = 3,  // total tiles of large groups
  CTR_NLARGE = 4,  // number of large groups
  CTR_NSEG = 5,    // number of groups in the current list
  CTR_ERR = 6,
  CTR_LELEMS = 7,  // total elements in large groups
  CTR_ROUND = 8,   // index of the current suffix-sort round (-1 before round 0)
  CTR_RAN = 9,     // rounds that had groups to sort
  CTR_RKN = 10,    // blocks whose ranks are built this round
  CTR_KEPT = 11,   // blocks that keep ranks (never decreases during a sort)
  CTR_RPK = 12,    // round 1: packs sorted by composite keys (listed in rpk)
  CTR_N = 16
};

// Most prefix-doubling rounds a suffix sort runs (prefix 4 * 2^40 > any block).
constexpr int kMaxSortRounds = 42;

// Prefix length the groups of round r >= 1 already share (4, 8, 16, ...,
// capped at 2^29, far beyond any block).
__host__ __device__ inline int32_t round_prefix(int r) { return r >= 28 ? (1 << 29) : (4 << (r - 1)); }

// Work-item loop of a persistent kernel (grid a multiple of 8): workgroup b
// runs on XCD b mod 8, and the workgroups of XCD x take the items of the
// x-th contiguous eighth of [0, n), so that neighbouring items (the tiles
// of one block) share an L2.
struct ItemLoop {
  int i, end, step;
  __device__ explicit ItemLoop(int n) {
    const int x = (int)(blockIdx.x & 7u), per = (int)(gridDim.x >> 3);
    i = (int)((int64_t)n * x / 8) + (int)(blockIdx.x >> 3);
    end = (int)((int64_t)n * (x + 1) / 8);
    step = per;
  }
};

// Persistent grid sizes of the suffix-sort kernels (workgroups; multiples
// of 8), from the device's CU count and each kernel's residency.
struct SortGrids {
  int light = 2048;   // plan, list, histogram, head-summary kernels
  int keys = 2048;
  int small = 1024;   // LDS pack sorts (32-bit composite keys)
  int small64 = 512;  // LDS pack sorts (64-bit composite keys)
  int scatter[3] = {1024, 512, 256};  // radix scatter, 256/512/1024 threads
  int heads = 1024;
  int apply = 1024;
};

// batch totals: blocks, raw bytes n, RLE1 bytes m, RLE2 symbols s, payload
// bits W, and the sum over blocks of m x suffix-sort rounds
constexpr int kTotals = 6;

// ---- launchers (each enqueues kernels on `s`, no host synchronisation)
void launch_block_init(const Dev& d, hipStream_t s);
void launch_crc_rle1(const Dev& d, hipStream_t s);   // CRC + RLE1 -> T, m, present
void launch_bwt_init(const Dev& d, hipStream_t s);   // SA = identity, first group list
// The grid sizes of the sort kernels for this device (occupancy queries).
SortGrids sort_grids(int device);
// The plan of the next suffix-sort round: totals of its group list, round
// counter advanced (the counters a host reads back to size the round).
void launch_bwt_plan(const Dev& d, hipStream_t s);
// One suffix-sort round.  Every kernel takes its work from the round's
// group list in device memory and does nothing when there is none.
// exact == null: the round plans itself and the kernels run on persistent
// grids, so rounds are enqueued (or captured in a graph) with no host
// readback, and a round after the last one does nothing.  exact = the
// counters read back after launch_bwt_plan: one workgroup per work item.
// first: round 0 (keys from the text, 32-bit keys); otherwise a later round
// (keys gathered from the ranks, late_bits-bit keys).  scatter_nt: threads
// per radix-scatter workgroup (256, 512 or 1024); pass_ev (may be null):
// one event before each radix pass and one after the last.
void launch_bwt_round(const Dev& d, const SortGrids& g, bool first, int late_bits, int scatter_nt,
                      hipStream_t s, const hipEvent_t* pass_ev, const int32_t* exact);
void launch_mtf_rle2(const Dev& d, hipStream_t s);   // MTF, orig_ptr, RLE2 -> SYM, s, freq
void launch_mtf(const Dev& d, hipStream_t s);        // L gather, MTF, orig_ptr, n_in_use
void launch_rle2(const Dev& d, hipStream_t s);       // RLE2 (needs tc_fnz, n_in_use)
void launch_tile_scan_c(const Dev& d, hipStream_t s);
// helpers for the per-stage entry points
void launch_presence(const Dev& d, hipStream_t s);   // st.present from T[0..m)
void launch_sa_attach(const Dev& d, const int32_t* sa, hipStream_t s);  // caller SA -> SA entries
void launch_sa_strip(const Dev& d, int32_t* sa, hipStream_t s);         // SA entries -> caller SA
void launch_fnz(const Dev& d, hipStream_t s);        // tc_fnz from MTF[0..m)
void launch_sym_hist(const Dev& d, hipStream_t s);   // freq from SYM[0..s)
void launch_check_codes(const Dev& d, int n_groups, hipStream_t s);  // st.err if SYM / sel out of range
void launch_huffman(const Dev& d, hipStream_t s);    // tables, selectors, codes, headers
void launch_pack(const Dev& d, hipStream_t s);       // payload words per block
void launch_assemble(const Dev& d, hipStream_t s);   // stream layout + placement into out
void launch_block_totals(const Dev& d, hipStream_t s);  // totals[0..kTotals)

}  // namespace bzg
