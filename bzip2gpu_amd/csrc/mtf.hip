// mtf.hip - forward MTF of the BWT last column, origPtr, and RLE2.
//
// MTF (reference: src/compression/fmtf.cu:40-97, :147-230).  The MTF list at
// any point is the set of present bytes ordered by most recent occurrence,
// never-seen bytes after them in ascending order (the initial list).  So the
// list state is a vector lastocc[256], states compose by element-wise max,
// and the state entering each 4096-symbol segment is an exclusive max-scan
// of per-segment last-occurrence vectors (k_mtf_scan).  Within a segment one
// wave handles 64 symbols per step: for lane i with symbol c,
//   c seen earlier in the step at lane p:  out = #distinct symbols in (p, i)
//   otherwise:  out = pos(c) + #{first occurrences j < i with pos(c_j) > pos(c)}
// where pos() is the list position entering the step.
//
// RLE2 (reference: src/compression/rle2.cu:11-77): nonzero v -> v+1, a zero
// run of length k -> the bits of k+1 below its leading one, LSB first
// (RUNA = 0, RUNB = 1), EOB = n_in_use + 1 appended.  Same count/scan/write
// structure as RLE1, with run ends from per-tile first-nonzero summaries.
#include "pipeline.h"

namespace bzg {

void launch_tile_scan_c(const Dev& d, hipStream_t s);

// L[i] = T[(SA[i]-1) mod m] (fmtf.cu:55) from the byte each SA entry carries,
// orig_ptr, per-tile last occurrence.
// Each wave walks 1024 consecutive rows 64 at a time; the last occurrence of
// every byte value is an LDS atomic max over the rows' positions (one
// counter set per wave, merged at the end).
__global__ void __launch_bounds__(256) k_mtf_gather(Dev d) {
  __shared__ int32_t lo[4][256];
  const int tid = threadIdx.x, wv = tid >> 6, lane = lane_id();
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int m = d.st[b].m;
  const int k = t - D.tile_c0;
  const int start = k * kTile;
  if (start >= m) return;
  lo[0][tid] = -1; lo[1][tid] = -1; lo[2][tid] = -1; lo[3][tid] = -1;
  __syncthreads();
  const int n = min(kTile, m - start);
  uint32_t sv[16];  // all 16 loads in flight before the first use
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = wv * 1024 + q * 64 + lane;
    sv[q] = e < n ? d.SA[D.off + start + e] : 0u;
  }
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = wv * 1024 + q * 64 + lane;
    const bool valid = e < n;
    const int i = start + e;
    uint32_t c = 0;
    if (valid) {
      const uint32_t v = sv[q];
      c = v >> 24;  // T[(start - 1) mod m], carried by the sort
      d.MTF[D.off + i] = (uint8_t)c;
      if ((v & kSAMask) == 0) d.st[b].orig_ptr = i;
    }
    if (valid) atomicMax(&lo[wv][c], i);
  }
  __syncthreads();
  d.lo[(int64_t)t * 256 + tid] = max(max(lo[0][tid], lo[1][tid]), max(lo[2][tid], lo[3][tid]));
}

// Per block: the set of byte values (those with a last occurrence in some
// segment; the last column is a permutation of the RLE1 text) and the
// entering list state of every segment (exclusive max-scan).
__global__ void __launch_bounds__(256) k_mtf_scan(Dev d) {
  __shared__ int32_t tmp[8];
  const int b = blockIdx.x;
  const BlockDesc D = d.desc[b];
  const int m = d.st[b].m;
  const int s = threadIdx.x;
  const int nt = tiles_for(m);
  bool present = false;
  for (int k0 = 0; k0 < nt; k0 += 16) {
    int v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = k0 + q < nt ? d.lo[(int64_t)(D.tile_c0 + k0 + q) * 256 + s] : -1;
#pragma unroll
    for (int q = 0; q < 16; q++) present |= v[q] >= 0;
  }
  const uint64_t pm = __ballot(present);
  if (lane_id() == 0) {
    d.st[b].present[2 * (s >> 6)] = (uint32_t)pm;
    d.st[b].present[2 * (s >> 6) + 1] = (uint32_t)(pm >> 32);
  }
  int tot;
  int rank = wg_excl_sum<256, int>(present ? 1 : 0, tmp, &tot);
  if (s == 0) d.st[b].n_in_use = tot;
  int E = present ? -1 - rank : (int)0x80000000;
  // 16 tiles' loads in flight per step (the scan itself is cheap)
  for (int k0 = 0; k0 < nt; k0 += 16) {
    int v[16];
#pragma unroll
    for (int q = 0; q < 16; q++)
      v[q] = k0 + q < nt ? d.lo[(int64_t)(D.tile_c0 + k0 + q) * 256 + s] : -1;  // -1: not in this segment
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (k0 + q < nt) d.lo[(int64_t)(D.tile_c0 + k0 + q) * 256 + s] = E;
      if (present && v[q] >= 0) E = v[q];  // later segments hold later positions
    }
  }
}

// Transpose of a 64x64 bit matrix held one row per lane (bit c of lane r is
// element (r, c)), as two 32-bit words per lane.  Stage s (32, 16, ..., 1)
// swaps the off-diagonal s x s blocks: lane r trades with lane r ^ s the bits
// whose column index differs from r in bit s.  The lane exchanges use
// gfx950's cross-lane VALU paths, no LDS: v_permlane32_swap (s = 32),
// v_permlane16_swap (s = 16), DPP row rotate (s = 8), DPP row shifts (s = 4)
// and DPP quad permutes (s = 2, 1).
namespace {

constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kDppRowShl4 = 0x104;
constexpr int kDppRowShr4 = 0x114;
constexpr int kDppRowRor8 = 0x128;

// word of lane ^ S (S < 16)
template <int S>
__device__ inline uint32_t xor_partner(uint32_t w) {
  if constexpr (S == 1) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)w, kDppQuadXor1, 0xF, 0xF, false);
  } else if constexpr (S == 2) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)w, kDppQuadXor2, 0xF, 0xF, false);
  } else if constexpr (S == 4) {
    // lanes 0-3 and 8-11 of each row read lane + 4, the others lane - 4
    const int up = __builtin_amdgcn_mov_dpp((int)w, kDppRowShl4, 0xF, 0x5, false);
    return (uint32_t)__builtin_amdgcn_update_dpp(up, (int)w, kDppRowShr4, 0xF, 0xA, false);
  } else {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)w, kDppRowRor8, 0xF, 0xF, false);
  }
}

// Stages s = 1, 2, 4 on one word.  A lane with bit s clear keeps the low s-bit
// halves of each 2s-bit field and takes the partner's low halves into the
// high ones; a lane with bit s set keeps the high halves and takes the
// partner's high halves into the low ones.
template <int S>
__device__ inline uint32_t tstage_bits(uint32_t w, bool clear) {
  constexpr uint32_t hi = S == 1 ? 0xAAAAAAAAu : S == 2 ? 0xCCCCCCCCu : 0xF0F0F0F0u;
  const uint32_t p = xor_partner<S>(w);
  const uint32_t t = __builtin_amdgcn_alignbit(p, p, clear ? 32 - S : S);  // rotate
  const uint32_t keep = clear ? ~hi : hi;
  return (w & keep) | (t & ~keep);
}

__device__ inline uint64_t bit_transpose64(uint64_t x) {
  const int lane = lane_id();
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  {  // s = 32: lanes 0-31 take the low words of lanes 32-63 as high words
    const auto r = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
    lo = r[0];
    hi = r[1];
  }
  {  // s = 16: 16-bit fields; after the swap a[0] holds the own word in rows 0, 2
     // and the partner's in rows 1, 3, a[1] the reverse
    const uint32_t sel = (lane & 16) ? 0x07060302u : 0x05040100u;
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    lo = __builtin_amdgcn_perm(a[1], a[0], sel);
    hi = __builtin_amdgcn_perm(b[1], b[0], sel);
  }
  {  // s = 8: bytes
    const uint32_t sel = (lane & 8) ? 0x03070105u : 0x06020400u;
    lo = __builtin_amdgcn_perm(xor_partner<8>(lo), lo, sel);
    hi = __builtin_amdgcn_perm(xor_partner<8>(hi), hi, sel);
  }
  lo = tstage_bits<4>(lo, !(lane & 4));
  hi = tstage_bits<4>(hi, !(lane & 4));
  lo = tstage_bits<2>(lo, !(lane & 2));
  hi = tstage_bits<2>(hi, !(lane & 2));
  lo = tstage_bits<1>(lo, !(lane & 1));
  hi = tstage_bits<1>(hi, !(lane & 1));
  return ((uint64_t)hi << 32) | lo;
}

}  // namespace

// The MTF workgroup is one wave, whose LDS accesses execute in program
// order: a barrier is only needed to keep the compiler from moving them.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave per segment: the MTF proper.  The segment is staged in LDS and
// the list is kept as s2p[] (list position of every present byte, one byte
// each).  After each 64-symbol step the K distinct bytes of the step take
// positions 0..K-1 by last occurrence, and every other present byte moves
// from position p to K + p - rank[p], rank[p] = #(step bytes whose old
// position is below p).
//
// A wave takes d.mtf_run consecutive capacity tiles: a segment that follows
// the previous one of the same block starts from the list state that one
// ended with, so the state entering it is computed only for a run's first
// segment of each block (256 x 256 compares).
__global__ void __launch_bounds__(64) k_mtf_apply(Dev d) {
  __shared__ int32_t E[256];
  __shared__ __align__(16) uint8_t buf[kTile];
  __shared__ __align__(4) uint8_t s2p[256];
  __shared__ __align__(4) uint8_t pflag[256];  // old positions of this step's bytes
  __shared__ __align__(4) uint8_t sflag[256];  // this step's bytes
  __shared__ __align__(4) uint8_t rank[256];
  __shared__ __align__(4) uint8_t newpos[256];
  __shared__ unsigned long long pm[256];  // per-byte lane masks of the current step
  const int lane = lane_id();
  const uint64_t below = lanemask_lt();
  const int w = (int)xcd_remap(blockIdx.x, gridDim.x);
  int held = -1;  // capacity tile whose exit list state s2p holds
  const int run = d.mtf_run;
  for (int j = 0; j < run; j++) {
    const int t = w * run + j;
    if (t >= d.ntile_c) break;
    const int b = d.tc_blk[t];
    const BlockDesc D = d.desc[b];
    const int m = d.st[b].m;
    const int k = t - D.tile_c0;
    const int start = k * kTile;
    if (start >= m) continue;
    const int n = min(kTile, m - start);
    uint8_t* mtf = d.MTF + D.off + start;  // 256-byte aligned segment
    wave_sync();  // the previous segment's copy-out has read buf
    for (int o = lane * 16; o < n; o += 1024) {
      if (o + 16 <= n) {
        *reinterpret_cast<uint4*>(buf + o) = *reinterpret_cast<const uint4*>(mtf + o);
      } else {
        for (int q = o; q < n; q++) buf[q] = mtf[q];
      }
    }
    // the previous segment of the same block left the list state this one
    // enters with (and the step flags cleared)
    if (held != t - 1 || k == 0) {
      for (int q = lane; q < 256; q += 64) {
        E[q] = d.lo[(int64_t)t * 256 + q];
        pflag[q] = 0;
        sflag[q] = 0;
        pm[q] = 0ull;
      }
      wave_sync();
      // list position of every present byte: # present bytes more recent
      {
        int e4[4], p4[4] = {0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; r++) e4[r] = E[4 * lane + r];
        for (int r = 0; r < 256; r++) {
          const int er = E[r];
#pragma unroll
          for (int q = 0; q < 4; q++) p4[q] += er > e4[q] ? 1 : 0;
        }
        reinterpret_cast<uint32_t*>(s2p)[lane] =
            (uint32_t)(p4[0] & 255) | ((uint32_t)(p4[1] & 255) << 8) | ((uint32_t)(p4[2] & 255) << 16) |
            ((uint32_t)(p4[3] & 255) << 24);
      }
    }
    wave_sync();
    int fnz = 0x7fffffff;
    for (int base = 0; base < n; base += 64) {
      const int i = base + lane;
      const bool valid = i < n;
      const uint32_t c = valid ? buf[i] : 0u;
      // lanes holding the same byte: each lane ORs its bit into its byte's
      // LDS mask and reads the mask back (cleared again below)
      if (valid) atomicOr(&pm[c], 1ull << lane);
      wave_sync();
      const uint64_t peers = valid ? pm[c] : 0ull;
      const uint64_t pb = peers & below;
      const bool has_prev = valid && pb != 0;
      const int prev = has_prev ? 63 - __clzll(pb) : -1;
      const uint64_t above = valid ? (peers & ~below & ~(1ull << lane)) : 0ull;
      const int nxt = above ? __builtin_ctzll(above) : 64;
      const int mypos = valid ? (int)s2p[c] : 0;
      const bool first = valid && !has_prev;
      // Repeated byte: distinct bytes strictly between its previous lane and
      // this lane = lanes j in (prev, lane) whose next occurrence is >= lane.
      // Row j of a 64x64 bit matrix marks the lanes i with j < i <= nxt_j;
      // its transpose gives, for lane i, the set of such j.
      uint64_t row = 0;
      if (valid) {
        uint64_t upto = nxt >= 63 ? ~0ull : ((2ull << nxt) - 1ull);
        row = upto & ~((2ull << lane) - 1ull);
        if (lane == 63) row = 0;
      }
      const uint64_t col = bit_transpose64(row);
      // First occurrence: earlier first occurrences whose old position is
      // behind this one, by a bit-serial compare of the positions (MSB first).
      const uint64_t fm = ballot(first);
      const int K = __popcll(fm);
      // gt: earlier first lanes with a larger position, eq: those equal so far
      // (MSB first); nm is all ones where this lane's bit is clear
      uint32_t gl = 0, gh = 0;
      uint32_t el = (uint32_t)(fm & below), eh = (uint32_t)((fm & below) >> 32);
#pragma unroll
      for (int bb = 7; bb >= 0; bb--) {
        const uint32_t nm = (uint32_t)__builtin_amdgcn_sbfe(~mypos, bb, 1);
        const uint64_t B = ballot(nm == 0) & fm;
        const uint32_t bl = (uint32_t)B, bh = (uint32_t)(B >> 32);
        gl |= el & bl & nm;
        gh |= eh & bh & nm;
        el &= bl ^ nm;
        eh &= bh ^ nm;
      }
      int out;
      if (has_prev) {
        const uint64_t between = below & ~((2ull << prev) - 1ull);
        out = __popcll(col & between);
      } else {
        out = mypos + __popc(gl) + __popc(gh);
      }
      if (valid) buf[i] = (uint8_t)out;
      const uint64_t nz = ballot(valid && out != 0);
      if (nz && fnz == 0x7fffffff) fnz = start + base + __builtin_ctzll(nz);
      // list update
      const bool last = valid && nxt == 64;
      const uint64_t lastm = ballot(last);
      if (first) {
        pflag[mypos] = 1;
        sflag[c] = 1;
      }
      if (last) newpos[c] = (uint8_t)(__popcll(lastm >> lane) - 1);
      if (valid) pm[c] = 0ull;  // every lane has read its mask (in program order)
      wave_sync();
      {
        // flags below this lane's four positions, from one ballot per byte slot
        const uint32_t f4 = reinterpret_cast<const uint32_t*>(pflag)[lane];
        const int ex = mbcnt(ballot(f4 & 1u)) + mbcnt(ballot((f4 >> 8) & 1u)) + mbcnt(ballot((f4 >> 16) & 1u)) +
                       mbcnt(ballot((f4 >> 24) != 0u));
        const int r1 = ex + (int)(f4 & 1u), r2 = r1 + (int)((f4 >> 8) & 1u), r3 = r2 + (int)((f4 >> 16) & 1u);
        reinterpret_cast<uint32_t*>(rank)[lane] =
            (uint32_t)ex | ((uint32_t)r1 << 8) | ((uint32_t)r2 << 16) | ((uint32_t)r3 << 24);
      }
      wave_sync();
      {
        const uint32_t sp4 = reinterpret_cast<const uint32_t*>(s2p)[lane];
        const uint32_t sf4 = reinterpret_cast<const uint32_t*>(sflag)[lane];
        const uint32_t nw4 = reinterpret_cast<const uint32_t*>(newpos)[lane];
        int rk[4];
#pragma unroll
        for (int r = 0; r < 4; r++) rk[r] = rank[(sp4 >> (8 * r)) & 255u];
        uint32_t np4 = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int op = (int)((sp4 >> (8 * r)) & 255u);
          const int np = ((sf4 >> (8 * r)) & 1u) ? (int)((nw4 >> (8 * r)) & 255u) : K + op - rk[r];
          np4 |= (uint32_t)(np & 255) << (8 * r);
        }
        wave_sync();
        reinterpret_cast<uint32_t*>(s2p)[lane] = np4;
        if (first) {
          pflag[mypos] = 0;
          sflag[c] = 0;
        }
      }
      wave_sync();
    }
    for (int o = lane * 16; o < n; o += 1024) {
      if (o + 16 <= n) {
        *reinterpret_cast<uint4*>(mtf + o) = *reinterpret_cast<const uint4*>(buf + o);
      } else {
        for (int q = o; q < n; q++) mtf[q] = buf[q];
      }
    }
    if (lane == 0) d.tc_fnz[t] = fnz;
    held = t;
  }
}

__device__ inline int ilog2(uint32_t v) { return 31 - __clz(v); }

// RLE2 count / write over capacity tiles, one wave per tile, four tiles per
// 256-thread workgroup; lane l holds MTF bytes [64 l, 64 l + 64) of its tile
// as 16 words.  Byte flags are words with 0x80 in each flagged byte.  A
// nonzero byte v emits the symbol v + 1; the first zero of a zero run of
// length L emits the ilog2(L + 1) bits of L + 1 below its leading one, low
// bit first (0 = RUNA, 1 = RUNB; src/compression/rle2.cu:47-67); other
// zeros emit nothing.
constexpr uint32_t kHi8 = 0x80808080u;
__device__ inline uint32_t nz80b(uint32_t x) { return (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & kHi8; }
__device__ inline uint32_t low_flags8(int r) { return r >= 4 ? kHi8 : (r <= 0 ? 0u : kHi8 & ((1u << (8 * r)) - 1u)); }

// Zero-run length of the run starting at byte q of word i: up to the next
// nonzero byte (NZ: nonzero flags; nx: lane-relative first nonzero after
// the word).
__device__ inline int zrun_len(uint32_t NZ, int nx, int i, int q) {
  const uint32_t after = q < 3 ? NZ >> (8 * q + 8) : 0u;
  const int e = after ? 4 * i + q + 1 + (__builtin_ctz(after) >> 3) : nx;
  return e - (4 * i + q);
}

// A wave stages its tile's symbols (at most one per byte) in LDS with one
// word of padding after every 64 symbols: lanes stage their own runs of
// symbols side by side, and without it the lanes' stores at equal offsets
// in their runs would fall on the same banks.
constexpr int kRle2Stage = kTile + 2 * (kTile / 64) + 8;
__device__ inline int spos(int j) { return j + 2 * (j >> 6); }

template <bool WRITE>
__device__ inline void rle2_wave(const Dev& d, uint16_t* ob, int32_t* hist) {
  const int tid = threadIdx.x, lane = lane_id();
  const int t = 4 * (int)xcd_remap(blockIdx.x, gridDim.x) + (tid >> 6);
  if (t >= d.ntile_c) return;
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int m = d.st[b].m;
  const int k = t - D.tile_c0;
  const int start = k * kTile;
  if (start >= m) {
    if (!WRITE && lane == 0) d.tc_cnt[t] = 0;
    return;
  }
  const int n = min(kTile, m - start);
  const int q0 = lane * 64;
  const int c = max(0, min(64, n - q0));
  // the tile's bytes (capacity slots are 256-byte aligned and readable past m)
  const uint8_t* src = d.MTF + D.off + start;
  uint32_t w[16];
  if (c > 0) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint4 v = *reinterpret_cast<const uint4*>(src + q0 + 16 * j);
      w[4 * j] = v.x; w[4 * j + 1] = v.y; w[4 * j + 2] = v.z; w[4 * j + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 16; i++) w[i] = 0u;
  }
  // first nonzero at or after the tile's end (block-relative), from the
  // later tiles' first nonzero positions, else m
  int carry = 0x7fffffff;
  for (int j = k + 1 + lane; j < tiles_for(m); j += 64) {
    const int v = d.tc_fnz[D.tile_c0 + j];
    if (v != 0x7fffffff) { carry = v; break; }
  }
  carry = wave_min(carry);
  if (carry == 0x7fffffff) carry = m;
  // nonzero flags; the byte before the tile counts as nonzero at the block
  // start, so that a zero there starts a run
  uint32_t NZ[16], ZS[16];
  bool prev_nz;
  {
    const uint32_t last = (uint32_t)__shfl_up((int)(w[15] >> 24), 1, 64);
    prev_nz = lane > 0 ? last != 0u : (start > 0 ? src[-1] != 0 : true);
  }
  uint32_t pf = prev_nz ? 0x80000000u : 0u;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t f = nz80b(w[i]);
    const uint32_t vm = low_flags8(c - 4 * i);
    // zero bytes whose previous byte is nonzero start a zero run
    ZS[i] = ~f & __builtin_amdgcn_alignbyte(f, pf, 3u) & vm;
    NZ[i] = f & vm;
    pf = f;
  }
  // first nonzero after each word (lane-relative; beyond the lane from the
  // later lanes, else the carry)
  int fl = 64;
#pragma unroll
  for (int i = 15; i >= 0; i--)
    if (NZ[i]) fl = 4 * i + (__builtin_ctz(NZ[i]) >> 3);
  int nxl = __shfl_down(wave_incl_min_rev(fl < 64 ? start + q0 + fl : 0x7fffffff), 1, 64);
  if (lane == 63) nxl = 0x7fffffff;
  nxl = min(nxl, carry) - (start + q0);
  int nx[16];
  {
    int nxt = nxl;
#pragma unroll
    for (int i = 15; i >= 0; i--) {
      nx[i] = nxt;
      if (NZ[i]) nxt = 4 * i + (__builtin_ctz(NZ[i]) >> 3);
    }
  }
  // symbols of the lane (a word holds at most two zero-run starts)
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    cnt += __popc(NZ[i]);
    uint32_t z = ZS[i];
#pragma unroll
    for (int r = 0; r < 2; r++) {
      if (z) {
        const int q = __builtin_ctz(z) >> 3;
        cnt += ilog2((uint32_t)zrun_len(NZ[i], nx[i], i, q) + 1u);
        z &= z - 1u;
      }
    }
  }
  if (!WRITE) {
    cnt = wave_sum(cnt);
    if (lane == 0) d.tc_cnt[t] = cnt;
    return;
  }
  const int incl = wave_incl_sum(cnt);
  int o = incl - cnt;
  const int total = __builtin_amdgcn_readlane(incl, 63);
  int n1 = 0, n0 = 0;  // RUNB / RUNA symbols of the lane
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t x = w[i], nzf = NZ[i];
    // the word's zero-run starts (at most two): byte, L + 1, symbol count
    int qa = 4, qb = 4, ca = 0, cb = 0;
    uint32_t ka = 0u, kb = 0u;
    uint32_t z = ZS[i];
    if (z) {
      qa = __builtin_ctz(z) >> 3;
      ka = (uint32_t)zrun_len(nzf, nx[i], i, qa) + 1u;
      ca = ilog2(ka);
      z &= z - 1u;
    }
    if (z) {
      qb = __builtin_ctz(z) >> 3;
      kb = (uint32_t)zrun_len(nzf, nx[i], i, qb) + 1u;
      cb = ilog2(kb);
    }
    // output order within the word: byte by byte, a nonzero byte's symbol
    // or a zero-run start's bits (start a precedes start b)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if ((nzf >> (8 * q + 7)) & 1u) {
        const uint32_t v = ((x >> (8 * q)) & 255u) + 1u;
        const int at = o + __popc(nzf & low_flags8(q)) + (qa < q ? ca : 0) + (qb < q ? cb : 0);
        ob[spos(at)] = (uint16_t)v;
        atomicAdd(&hist[v], 1);
      }
    }
    // the run bits of both starts in one loop
    {
      const int pa = o + __popc(nzf & low_flags8(qa));
      const int pb = o + __popc(nzf & low_flags8(qb)) + ca;
#pragma unroll 1
      for (int r = 0; r < ca + cb; r++) {
        const bool in_a = r < ca;
        const int at = in_a ? pa + r : pb + (r - ca);
        const uint32_t bit = in_a ? (ka >> r) & 1u : (kb >> (r - ca)) & 1u;
        ob[spos(at)] = (uint16_t)bit;
      }
    }
    o += __popc(nzf) + ca + cb;
    asm volatile("" : "+v"(o));  // one running offset (not a sum over earlier words at each use)
    const int ones = (ca ? __popc(ka) - 1 : 0) + (cb ? __popc(kb) - 1 : 0);
    n1 += ones;
    n0 += ca + cb - ones;
  }
  n1 = wave_sum(n1);
  n0 = wave_sum(n0);
  if (lane == 0) {
    if (n0) atomicAdd(&hist[0], n0);
    if (n1) atomicAdd(&hist[1], n1);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // copy-out as 32-bit words of two symbols (single symbols at the edges)
  uint16_t* dst = d.SYM + D.off + d.tc_off[t];
  const int lead = (int)((reinterpret_cast<uintptr_t>(dst) >> 1) & 1u);
  uint32_t* A0 = reinterpret_cast<uint32_t*>(reinterpret_cast<uintptr_t>(dst) & ~(uintptr_t)3);
  const int nw = (lead + total + 1) >> 1;
  const uint32_t* o32 = reinterpret_cast<const uint32_t*>(ob);
  for (int wi = lane; wi < nw; wi += 64) {
    const int oo = 2 * wi - lead;  // symbol index of the word's low half
    if (oo >= 0 && oo + 2 <= total) {
      // symbols oo, oo + 1: one staged word, or the halves of two (odd oo)
      const uint32_t v = lead ? __builtin_amdgcn_alignbyte(o32[spos(oo + 1) >> 1], o32[spos(oo - 1) >> 1], 2u)
                              : o32[spos(oo) >> 1];
      A0[wi] = v;
    } else {
      if (oo >= 0 && oo < total) dst[oo] = ob[spos(oo)];
      if (oo + 1 >= 0 && oo + 1 < total) dst[oo + 1] = ob[spos(oo + 1)];
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int q = lane; q < 260; q += 64) {
    const int h = hist[q];
    if (h) atomicAdd(&d.freq[(int64_t)b * 260 + q], h);
  }
}

__global__ void __launch_bounds__(256) k_rle2_count(Dev d) { rle2_wave<false>(d, nullptr, nullptr); }

__global__ void __launch_bounds__(256) k_rle2_write(Dev d) {
  __shared__ __align__(16) uint16_t ob[4][kRle2Stage];
  __shared__ int32_t hist[4][260];
  const int wv = threadIdx.x >> 6;
  for (int q = lane_id(); q < 260; q += 64) hist[wv][q] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  rle2_wave<true>(d, ob[wv], hist[wv]);
}

void launch_mtf(const Dev& d, hipStream_t s) {
  k_mtf_gather<<<d.ntile_c, 256, 0, s>>>(d);
  k_mtf_scan<<<d.nblk, 256, 0, s>>>(d);
  probe_begin("k_mtf_apply", s);
  // d.mtf_run is 1, 2, 4 or 8 (set by the engine)
  k_mtf_apply<<<(d.ntile_c + d.mtf_run - 1) / d.mtf_run, 64, 0, s>>>(d);
  probe_end("k_mtf_apply", s);
}

void launch_rle2(const Dev& d, hipStream_t s) {
  const int wgs = (d.ntile_c + 3) / 4;  // four tiles per workgroup, one per wave
  k_rle2_count<<<wgs, 256, 0, s>>>(d);
  launch_tile_scan_c(d, s);
  k_rle2_write<<<wgs, 256, 0, s>>>(d);
}

void launch_mtf_rle2(const Dev& d, hipStream_t s) {
  launch_mtf(d, s);
  launch_rle2(d, s);
}

}  // namespace bzg
