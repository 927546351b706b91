// common.h - shared device-side definitions for the bzip2 block pipeline on
// CDNA4 (gfx950, wave64).  All arithmetic in this library is integer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bzg {

constexpr int kWave = 64;
constexpr int kWG = 256;              // threads per workgroup for tile kernels
constexpr int kTile = 4096;           // bytes/elements per tile
constexpr int kTileItems = kTile / kWG;  // 16 per thread
constexpr int kMaxGroups = 6;
constexpr int kMaxAlpha = 258;
constexpr int kGSize = 50;            // symbols per selector group
constexpr int kMaxBlock = 900000;     // raw bytes per block (reference chunk size)
constexpr int kStageMax = 1 << 24;    // largest single block of a per-stage call
// Suffix-array entries hold the start index in the low 24 bits and the
// rotation's preceding byte (its BWT output byte) in the top 8 bits.
constexpr uint32_t kSAMask = 0x00FFFFFFu;
static_assert(kStageMax <= (int)kSAMask + 1, "start indices must fit 24 bits");
// Rank updates of a round are bucketed by start index, kRkBucket per bucket.
constexpr int kRkBucket = 8192;
constexpr uint32_t kCrcPoly = 0x04C11DB7u;

// Capacity (elements) reserved for a block of n raw bytes: RLE1 can turn
// every 4 raw bytes into 5 (a run of exactly four), so m <= 5*ceil(n/4).
__host__ __device__ inline int32_t cap_for(int32_t n) {
  int64_t c = 5 * ((int64_t)n + 3) / 4 + 64;
  return (int32_t)((c + 255) & ~255ll);
}
__host__ __device__ inline int32_t tiles_for(int32_t len) { return (len + kTile - 1) / kTile; }

// Static description of one block of a batch (host-built, read-only on device).
struct BlockDesc {
  uint64_t in_off;     // raw bytes offset into the batch input
  uint64_t off;        // element offset into cap-sized per-block arrays
  uint64_t sel_off;    // offset into the selector arrays (cap/50 + 2 per block)
  uint64_t w_off;      // offset (u32 words) into payload / header word arrays
  int32_t n;           // raw length (1..)
  int32_t cap;         // capacity, cap_for(n)
  int32_t tile_n0;     // first tile index of this block over raw bytes
  int32_t tile_c0;     // first tile index of this block over capacity
  int32_t stream;      // output stream index this block belongs to
  int32_t level;       // header level digit of its stream (1..9)
};

// Dynamic per-block results (device-written).
struct BlockState {
  uint32_t crc;
  int32_t m;           // RLE1 length
  int32_t s;           // RLE2 symbols incl. EOB
  int32_t orig_ptr;
  int32_t n_in_use;
  int32_t n_groups;
  int32_t n_sel;
  int32_t rounds;      // prefix-doubling rounds run
  uint32_t present[8]; // 256 presence bits of the RLE1 bytes
  int64_t bits;        // payload bits W
  int64_t hdr_bits;    // block header bits (magic .. coding tables)
  int64_t bit_off;     // bit offset of the block inside its stream
  int32_t err;         // nonzero if a device-side invariant failed
  int32_t rk;          // suffix-sort ranks: 0 not kept, 1 to be built this round, 2 kept
  int32_t rk_rows;     // suffix sort, round 1: rows in packs sorted by composite keys
};

// Per-output-stream layout (device-written by the layout kernel).  A whole
// stream (lead < 0) has the 4-byte header, its blocks and the end-of-stream
// trailer.  A segment of a stream written over several pipeline runs
// (lead = 0..7) has only its blocks' bits, starting `lead` bits into its
// first byte (the bits the earlier segments left in their last byte), and
// carries the combined CRC on from crc_in.
struct StreamState {
  int64_t bits;        // total bits incl. stream header/trailer (or lead), before pad
  int64_t byte_off;    // 4-byte aligned offset into the output buffer
  int64_t bytes;       // stream length in bytes
  uint32_t crc;        // combined stream CRC (after this segment's blocks)
  int32_t first_blk;   // first block index (blocks of a stream are consecutive)
  int32_t n_blk;
  int32_t level;
  int32_t lead;        // -1: whole stream; 0..7: segment starting at this bit
  uint32_t crc_in;     // combined CRC of the blocks before a segment
};

// ------------------------------------------------------------- lane helpers
__device__ inline int lane_id() { return __lane_id(); }

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

// The lanes among `lanes` whose low NBITS key bits equal this lane's.  Per
// bit: nb is all ones where the bit is set and m the lanes where it is set,
// so peers & ~(m ^ nb) keeps the lanes that agree on it, one three-input
// logic op per 32-bit half (truth table 0x90 is src0 & ~(src1 ^ src2)).  The
// empty asm keeps nb opaque, so that the ballot compares it instead of
// testing the key's bit a second time.
template <int NBITS>
__device__ inline uint64_t wave_match_bits(uint32_t key, uint64_t lanes) {
  uint32_t plo = (uint32_t)lanes, phi = (uint32_t)(lanes >> 32);
#pragma unroll
  for (int b = 0; b < NBITS; b++) {
    uint32_t nb = (uint32_t)__builtin_amdgcn_sbfe((int)key, b, 1);
    asm volatile("" : "+v"(nb));
    const uint64_t m = __ballot(nb != 0u);
    plo = __builtin_amdgcn_bitop3_b32(plo, (uint32_t)m, nb, 0x90);
    phi = __builtin_amdgcn_bitop3_b32(phi, (uint32_t)(m >> 32), nb, 0x90);
  }
  return ((uint64_t)phi << 32) | plo;
}

// number of set bits of m in lanes below this lane
__device__ inline int mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive wave scans (shuffle based).
template <typename T>
__device__ inline T wave_incl_sum(T v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T u = __shfl_up(v, d, 64);
    if (lane_id() >= d) v += u;
  }
  return v;
}
template <typename T>
__device__ inline T wave_incl_max(T v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T u = __shfl_up(v, d, 64);
    if (lane_id() >= d) v = u > v ? u : v;
  }
  return v;
}
template <typename T>
__device__ inline T wave_incl_min_rev(T v) {  // suffix min (lane..63)
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T u = __shfl_down(v, d, 64);
    if (lane_id() + d < 64) v = u < v ? u : v;
  }
  return v;
}
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
template <typename T>
__device__ inline T wave_xor(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v ^= __shfl_xor(v, d, 64);
  return v;
}
template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    T u = __shfl_xor(v, d, 64);
    v = u > v ? u : v;
  }
  return v;
}
template <typename T>
__device__ inline T wave_min(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    T u = __shfl_xor(v, d, 64);
    v = u < v ? u : v;
  }
  return v;
}

// Workgroup exclusive sum for blockDim.x == NT (multiple of 64).  `tmp`
// must hold NT/64 + 1 elements of T in LDS.  Returns exclusive prefix and
// writes the total to *total.
template <int NT, typename T>
__device__ inline T wg_excl_sum(T v, T* tmp, T* total) {
  int w = threadIdx.x >> 6, l = lane_id();
  T inc = wave_incl_sum(v);
  if (l == 63) tmp[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int i = 0; i < NT / 64; i++) { T t = tmp[i]; tmp[i] = run; run += t; }
    tmp[NT / 64] = run;
  }
  __syncthreads();
  T r = tmp[w] + inc - v;
  if (total) *total = tmp[NT / 64];
  __syncthreads();
  return r;
}

// Workgroup inclusive max-scan.
template <int NT, typename T>
__device__ inline T wg_incl_max(T v, T* tmp, T init) {
  int w = threadIdx.x >> 6, l = lane_id();
  T inc = wave_incl_max(v);
  if (l == 63) tmp[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = init;
    for (int i = 0; i < NT / 64; i++) { T t = tmp[i]; tmp[i] = run; run = t > run ? t : run; }
  }
  __syncthreads();
  T p = tmp[w];
  T r = inc > p ? inc : p;
  __syncthreads();
  return r;
}

// Workgroup suffix (reverse inclusive) min-scan.
template <int NT, typename T>
__device__ inline T wg_rev_incl_min(T v, T* tmp, T init) {
  int w = threadIdx.x >> 6, l = lane_id();
  T inc = wave_incl_min_rev(v);
  if (l == 0) tmp[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = init;
    for (int i = NT / 64 - 1; i >= 0; i--) { T t = tmp[i]; tmp[i] = run; run = t < run ? t : run; }
  }
  __syncthreads();
  T p = tmp[w];
  T r = inc < p ? inc : p;
  __syncthreads();
  return r;
}

// ------------------------------------------------------------ CRC helpers
// GF(2) product a*b mod P for MSB-first CRC registers (bit 31 = x^31).
__host__ __device__ constexpr inline uint32_t gf2_mulmod(uint32_t a, uint32_t b) {
  uint32_t r = 0;
  for (int i = 31; i >= 0; i--) {
    r = (r << 1) ^ ((r & 0x80000000u) ? kCrcPoly : 0u);
    if ((b >> i) & 1u) r ^= a;
  }
  return r;
}
// x^(8*len) mod P
__host__ __device__ constexpr inline uint32_t gf2_xpow8(uint64_t len) {
  uint32_t result = 1u;                 // polynomial "1"
  uint32_t base = 0x100u;               // x^8
  while (len) {
    if (len & 1) result = gf2_mulmod(result, base);
    base = gf2_mulmod(base, base);
    len >>= 1;
  }
  return result;
}

struct CrcTable {
  uint32_t t[256];
  constexpr CrcTable() : t() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i << 24;
      for (int j = 0; j < 8; j++) c = (c & 0x80000000u) ? (c << 1) ^ kCrcPoly : (c << 1);
      t[i] = c;
    }
  }
};

// XCD-aware remap: the dispatcher deals workgroups round-robin over the 8
// XCDs; give each XCD a contiguous range of logical tiles so that a block's
// tiles share one L2.
__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t n) {
  const uint32_t X = 8;
  uint32_t per = n / X, rem = n % X;
  uint32_t x = bid % X, k = bid / X;
  // XCD x owns logical range [x*per + min(x,rem), ...) of length per + (x<rem)
  uint32_t base = x * per + (x < rem ? x : rem);
  uint32_t len = per + (x < rem ? 1 : 0);
  return k < len ? base + k : bid;  // bid >= X*len only when n is tiny
}

}  // namespace bzg
