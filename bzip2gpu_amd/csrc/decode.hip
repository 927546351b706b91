// decode.hip - device bzip2 decoder: block discovery, Huffman decoding,
// inverse RLE2 + MTF, inverse BWT, inverse RLE1 and CRC checks.
//
// The wire format is the one the reference's stream driver writes
// (SebastianBoteroLeonik/bzip2gpu src/compression/compress.cu:184-375 and
// :284-375 for the block fields) and stock bzip2 reads; the reference itself
// has no decoder (TODO.md:28-34).  Every stage undoes one compression stage:
//
//   1. k_dec_words / k_dec_magic: the input as big-endian 32-bit words, and
//      every bit offset holding the 48-bit block or end-of-stream magic.
//      Candidates inside compressed data are possible; only those on the
//      chain that starts after the stream header and continues at each
//      block's end are used (the host walks it).
//   2. k_dec_huff: one wave per candidate parses the block header (CRC,
//      origPtr, symbol map, selectors, delta-coded lengths), builds 10-bit
//      lookup tables per coding table and decodes the symbols: every step
//      the 64 lanes decode a codeword at the 64 bit offsets after the
//      current position, and the chain of real codeword starts is followed
//      with lane reads, so one step yields all symbols starting in those 64
//      bits (up to the end of the 50-symbol group).
//   3. k_dec_mtf1 / k_dec_mtf_scan / k_dec_mtf2: inverse RLE2 and MTF.  A
//      tile of 4096 symbols is replayed from the identity list, which gives
//      the list slot every symbol outputs and the tile's permutation; the
//      per-block scan composes the permutations into each tile's starting
//      list, and the write pass expands zero runs and maps slots to bytes
//      (the L column).
//   4. k_dec_hist / k_dec_colscan / k_dec_scatter: a stable counting sort of
//      L by byte gives the FL map: entry q = (row of the q-th smallest
//      L byte << 8) | that byte.
//   5. k_dec_walk1 / k_dec_rank / k_dec_copy / k_dec_walk2 / k_dec_repl: the
//      RLE1 bytes are read off the FL map's cycle through origPtr.  Rows
//      that are multiples of 256 (and origPtr) are splitters; a walker per
//      splitter follows the map to the next splitter, keeping the bytes it
//      passes (up to 1024) in its staging slot; the per-block ranking
//      (pointer jumping in LDS) gives each splitter on origPtr's cycle its
//      position, the copy pass moves the staged bytes there, and the rare
//      longer stretches are finished by a second walk from where the first
//      stopped keeping bytes.  A block
//      that is a string repeated k times has k cycles of m/k rows; like
//      stock bzip2, which follows origPtr's cycle for m steps, the output is
//      that cycle repeated (k_dec_repl).
//   6. k_dec_rle1a / k_dec_rle1scan / k_dec_rle1b / k_dec_crc: inverse RLE1
//      as a 5-state machine (state = length of the current run of equal
//      bytes, 4 = next byte is a repeat count); tiles compose their
//      transfer functions, the write pass expands runs and computes the
//      tile's CRC register, and the per-block pass checks the block CRC.
#include "crc.h"
#include "decode.h"
#include "pipeline.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace bzg {

namespace {

constexpr uint64_t kBlockMagic = 0x314159265359ull;
constexpr uint64_t kEosMagic = 0x177245385090ull;
constexpr int kMaxLen = 20;
constexpr int kPadWords = 16;
constexpr int kRing = 256;       // LDS ring of input words for the symbol loop

// ---------------------------------------------------------------- bit input

__device__ inline uint32_t bits32_at(const uint32_t* W, int64_t pos) {
  const uint64_t w = (uint64_t)pos >> 5;
  const uint32_t sh = (uint32_t)pos & 31u;
  const uint64_t x = ((uint64_t)W[w] << 32) | W[w + 1];
  return (uint32_t)(x >> (32 - sh));
}

// Uniform MSB-first bit reader; `fetch(i)` returns input word i.
template <class F>
struct Bits {
  F fetch;
  uint64_t buf;  // next bits, MSB first
  int avail;     // valid bits in buf
  uint64_t wi;   // next word to load
  int64_t pos;   // bit position of buf's MSB
  __device__ Bits(F f, int64_t p) : fetch(f) {
    wi = (uint64_t)p >> 5;
    const int sh = (int)(p & 31);
    buf = (((uint64_t)fetch(wi) << 32) | fetch(wi + 1)) << sh;
    avail = 64 - sh;
    wi += 2;
    pos = p;
  }
  __device__ void fill() {
    if (avail <= 32) {
      buf |= (uint64_t)fetch(wi) << (32 - avail);
      wi++;
      avail += 32;
    }
  }
  __device__ uint32_t peek(int n) const { return (uint32_t)(buf >> (64 - n)); }
  __device__ void skip(int n) {
    buf <<= n;
    avail -= n;
    pos += n;
  }
  __device__ uint32_t get(int n) {
    fill();
    const uint32_t v = peek(n);
    skip(n);
    return v;
  }
};

// ---------------------------------------------------------------- stage 1

__global__ void __launch_bounds__(256) k_dec_words(const uint8_t* __restrict__ in, int64_t len,
                                                   uint32_t* __restrict__ W, int64_t nwords) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nwords) return;
  const int64_t b = i * 4;
  uint32_t v = 0;
  if (b + 4 <= len) {
    v = ((uint32_t)in[b] << 24) | ((uint32_t)in[b + 1] << 16) | ((uint32_t)in[b + 2] << 8) | in[b + 3];
  } else {
    for (int k = 0; k < 4; k++) v = (v << 8) | (b + k < len ? (uint32_t)in[b + k] : 0u);
  }
  W[i] = v;
}

// One thread per word: the 32 bit offsets starting in it.
__global__ void __launch_bounds__(256) k_dec_magic(const uint32_t* __restrict__ W, int64_t nbits,
                                                   DecCand* __restrict__ cand, int32_t cap,
                                                   int32_t* __restrict__ count) {
  const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t p0 = w * 32;
  if (p0 + 48 > nbits) return;
  const uint64_t x = ((uint64_t)W[w] << 32) | W[w + 1];
  const uint64_t z = W[w + 2];
  for (int k = 0; k < 32; k++) {
    const uint64_t v = ((x << k) | (k ? (z >> (32 - k)) : 0ull)) >> 16;
    const int kind = v == kBlockMagic ? 0 : (v == kEosMagic ? 1 : -1);
    if (kind >= 0 && p0 + k + 48 <= nbits) {
      const int i = atomicAdd(count, 1);
      if (i < cap) {
        cand[i].bit = p0 + k;
        cand[i].kind = kind;
        cand[i].sym_cap = 0;
        cand[i].sym_off = 0;
      }
    }
  }
}

// ---------------------------------------------------------------- stage 2

// Lookup tables of LB bits (codes up to LB bits in one read): LB = 12 takes
// about 56 KB of LDS (two blocks' decodes per CU), LB = 9 about 14 KB
// (eleven per CU; the selectors live in global memory).  The first kHdr
// words of a block (its header) are read from LDS, later ones from memory.
template <int LB>
struct HuffLds {
  static constexpr int kLong = kMaxLen - LB;  // code lengths past the table
  static constexpr int kHdr = LB >= 12 ? 2048 : 1024;
  union {
    uint16_t lut[kMaxGroups * (1 << LB)];  // (symbol << 5) | length, 0 = none
    uint32_t hw[kHdr];                     // header words (before the tables exist)
  };
  uint32_t ring[kRing];
  uint16_t perm[kMaxGroups * kMaxAlpha];       // symbols by (length, symbol)
  uint8_t lens[kMaxGroups * kMaxAlpha];
  int32_t first[kMaxGroups * 21], cnt[kMaxGroups * 21], base[kMaxGroups * 21];
  int32_t maxlen[kMaxGroups];
  // codes longer than the lookup table (lengths LB+1..20): left-aligned
  // 20-bit limits (first code + count) and perm index minus first code, per
  // length (padded to whole 16-byte loads)
  uint32_t lim[kMaxGroups][(kLong + 3) / 4 * 4] __attribute__((aligned(16)));
  int32_t dlt[kMaxGroups][kLong];
  int32_t run[32];
  uint8_t seq[256];
};

// Codes longer than the lookup table: the length is the smallest l whose
// left-aligned limit exceeds the 20 peeked bits (empty lengths repeat the
// previous limit, so the smallest such l is a used length).
template <int LB>
__device__ inline uint32_t long_code(const HuffLds<LB>& H, int t, uint32_t p20) {
  constexpr int NL = HuffLds<LB>::kLong;
  uint32_t lim[(NL + 3) / 4 * 4];
#pragma unroll
  for (int q = 0; q < (NL + 3) / 4; q++)
    *reinterpret_cast<uint4*>(&lim[4 * q]) = *reinterpret_cast<const uint4*>(&H.lim[t][4 * q]);
  int l = 0;
#pragma unroll
  for (int q = NL - 1; q >= 0; q--) l = p20 < lim[q] ? LB + 1 + q : l;
  if (l == 0) return 0u;
  const int i = (int)(p20 >> (kMaxLen - l)) + H.dlt[t][l - LB - 1];
  return ((uint32_t)H.perm[t * kMaxAlpha + i] << 5) | (uint32_t)l;
}

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// One 64-thread workgroup per candidate (idx[blockIdx.x] indexes cands).
template <int kLutBits>
__global__ void __launch_bounds__(64) k_dec_huff(const uint32_t* __restrict__ W, int64_t nbits, int64_t nwords,
                                                 const DecCand* __restrict__ cands,
                                                 const int32_t* __restrict__ idx, DecRes* __restrict__ res,
                                                 uint16_t* __restrict__ sym, uint8_t* __restrict__ selbuf) {
  __shared__ HuffLds<kLutBits> H;
  const int lane = threadIdx.x;
  const int ci = idx[blockIdx.x];
  const DecCand c = cands[ci];
  // this candidate's selectors: raw table indexes, then the MTF-decoded ones
  // (a separate area, so no lane reads a line it read before another lane
  // wrote it)
  uint8_t* sel_raw = selbuf + (int64_t)ci * (2 * kDecSelCap);
  uint8_t* sel = sel_raw + kDecSelCap;
  DecRes* R = res + ci;
  const uint64_t wlast = (uint64_t)nwords - 1;  // a zero pad word
  auto gw = [&](uint64_t i) -> uint32_t { return W[i < wlast ? i : wlast]; };
  if (c.kind == 1) {
    Bits<decltype(gw)> br(gw, c.bit + 48);
    const uint32_t crc = br.get(32);
    const int64_t end = c.bit + 80;
    const int64_t nbytes = nbits >> 3;
    const int64_t nb = (end + 7) >> 3;
    if (lane == 0) {
      R->end_bit = end;
      R->s = 0;
      R->status = end <= nbits ? kDecOk : kDecTruncated;
      R->crc = crc;
      R->orig_ptr = 0;
      R->n_in_use = 0;
      R->rest = nbytes - nb;
      R->next4 = nb < nbytes ? bits32_at(W, nb * 8) : 0u;
    }
    return;
  }
  // the words of the block header, read from LDS by the bit reader
  const uint64_t hb = (uint64_t)(c.bit + 48) >> 5;
  constexpr int kHdrWords = HuffLds<kLutBits>::kHdr;
  for (int i = lane; i < kHdrWords; i += 64) H.hw[i] = gw(hb + i);
  __syncthreads();
  auto hwf = [&](uint64_t i) -> uint32_t {
    const uint64_t d = i - hb;
    return d < (uint64_t)kHdrWords ? (uint32_t)uni((int)H.hw[d]) : gw(i);
  };
  Bits<decltype(hwf)> br(hwf, c.bit + 48);
  int status = kDecOk;
  const uint32_t bcrc = br.get(32);
  const uint32_t rnd = br.get(1);
  const int32_t orig = (int32_t)br.get(24);
  const uint32_t used = br.get(16);
  int nuse = 0;
  for (int i = 0; i < 16; i++) {
    if ((used >> (15 - i)) & 1u) {
      const uint32_t u = br.get(16);
      const bool on = lane < 16 && ((u >> (15 - lane)) & 1u);
      const uint64_t bm = ballot(on);
      if (on) H.seq[nuse + mbcnt(bm)] = (uint8_t)(i * 16 + lane);
      nuse += __popcll(bm);
    }
  }
  const int alpha = nuse + 2;
  const int ng = (int)br.get(3);
  const int nsel = (int)br.get(15);
  if (rnd) status = kDecRandomised;
  else if (nuse == 0 || ng < 2 || ng > kMaxGroups || nsel < 1) status = kDecBad;
  // selectors: unary table indexes (j ones, then a zero), 32 bits per step:
  // each zero bit ends a selector whose index is the run of ones before it
  if (status == kDecOk) {
    int k = 0, run = 0;
    bool bad = false;
    while (k < nsel) {
      br.fill();
      const uint32_t v = br.peek(32);
      const bool zero = lane < 32 && !((v >> (31 - lane)) & 1u);
      const uint64_t zm = ballot(zero);
      const int nz = __popcll(zm);
      const int take = min(nz, nsel - k);  // selectors ending in these 32 bits
      const int r = mbcnt(zm);
      if (zero && r < take) {
        const uint64_t below = zm & lanemask_lt();
        const int j = below ? lane - (63 - __clzll(below)) - 1 : run + lane;
        if (j >= ng) bad = true;
        if (k + r < kDecSelCap) sel_raw[k + r] = (uint8_t)j;
      }
      if (ballot(bad)) {
        status = kDecBad;
        break;
      }
      if (k + nz >= nsel) {  // the last selector ends at zero number take-1
        const int p = __ffsll((unsigned long long)ballot(zero && r == take - 1)) - 1;
        br.skip(p + 1);
        break;
      }
      run = nz ? 31 - (63 - __clzll(zm)) : run + 32;  // ones after the last zero
      if (run >= ng) {
        status = kDecBad;
        break;
      }
      br.skip(32);
      k += nz;
    }
  }
  // move-to-front over the table ids, 64 selectors per batch
  __syncthreads();  // the raw selectors are in memory
  if (status == kDecOk) {
    const int nk = min(nsel, kDecSelCap);
    uint32_t mtf = 0x543210u;
    for (int k0 = 0; k0 < nk; k0 += 64) {
      const int i = k0 + lane;
      const uint32_t raw = i < nk ? sel_raw[i] : 0u;
      uint32_t res = 0;
      const int cnt = min(64, nk - k0);
      for (int q = 0; q < cnt; q++) {
        const uint32_t j = (uint32_t)__builtin_amdgcn_readlane((int)raw, q);
        const uint32_t v = (mtf >> (4 * j)) & 15u;
        const uint32_t low = (1u << (4 * j)) - 1u;
        mtf = (mtf & ~((low << 4) | 15u)) | ((mtf & low) << 4) | v;
        res = lane == q ? v : res;
      }
      if (i < nk) sel[i] = (uint8_t)res;
    }
  }
  // code lengths: 5-bit start, then per symbol "1x" steps (x=0: +1, 1: -1) ending in "0"
  if (status == kDecOk) {
    for (int t = 0; t < ng && status == kDecOk; t++) {
      int cur = (int)br.get(5);
      for (int i = 0; i < alpha; i++) {
        for (;;) {
          if (cur < 1 || cur > kMaxLen) {
            status = kDecBad;
            break;
          }
          br.fill();
          const uint32_t b2 = br.peek(2);
          if (!(b2 & 2u)) {
            br.skip(1);
            break;
          }
          cur += (b2 & 1u) ? -1 : 1;
          br.skip(2);
        }
        if (status != kDecOk) break;
        if (lane == 0) H.lens[t * kMaxAlpha + i] = (uint8_t)cur;
      }
    }
  }
  if (br.pos > nbits && status == kDecOk) status = kDecTruncated;
  if (status == kDecOk) {
    // canonical codes (lengths ascending, symbols ascending within a length)
    __syncthreads();  // the header words are no longer read
    for (int i = lane; i < kMaxGroups * (1 << kLutBits); i += 64) H.lut[i] = 0;
    for (int i = lane; i < kMaxGroups * 21; i += 64) H.cnt[i] = 0;
    __syncthreads();
    for (int t = 0; t < ng; t++)
      for (int i = lane; i < alpha; i += 64) atomicAdd(&H.cnt[t * 21 + H.lens[t * kMaxAlpha + i]], 1);
    __syncthreads();
    bool over = false;
    if (lane < ng) {
      const int t = lane;
      int code = 0, ix = 0, ml = 0;
      for (int l = 1; l <= kMaxLen; l++) {
        const int k = H.cnt[t * 21 + l];
        H.first[t * 21 + l] = code;
        H.base[t * 21 + l] = ix;
        if (l > kLutBits) {
          H.lim[t][l - kLutBits - 1] = (uint32_t)(code + k) << (kMaxLen - l);
          H.dlt[t][l - kLutBits - 1] = ix - code;
        }
        if (k) ml = l;
        code += k;
        ix += k;
        if (code > (1 << l)) over = true;
        code <<= 1;
      }
      H.maxlen[t] = ml;
    }
    if (ballot(over)) status = kDecBad;
    __syncthreads();
  }
  if (status == kDecOk) {
    for (int t = 0; t < ng; t++) {
      if (lane <= kMaxLen) H.run[lane] = 0;
      __syncthreads();
      for (int i0 = 0; i0 < alpha; i0 += 64) {
        const int i = i0 + lane;
        const bool v = i < alpha;
        const int l = v ? H.lens[t * kMaxAlpha + i] : 0;
        const uint64_t peers = wave_match_bits<5>((uint32_t)l, ballot(v));
        const int r = __popcll(peers & lanemask_lt());
        const int before = v ? H.run[l] : 0;
        __syncthreads();
        if (v && r == 0) H.run[l] = before + __popcll(peers);
        if (v) {
          const int k = before + r;
          H.perm[t * kMaxAlpha + H.base[t * 21 + l] + k] = (uint16_t)i;
          if (l <= kLutBits) {
            const int code = H.first[t * 21 + l] + k;
            const int lo = code << (kLutBits - l), hi = (code + 1) << (kLutBits - l);
            for (int e = lo; e < hi; e++) H.lut[t * (1 << kLutBits) + e] = (uint16_t)((i << 5) | l);
          }
        }
        __syncthreads();
      }
    }
  }
  // symbols, one 50-symbol group per step: lane j decodes a codeword at each
  // of the 8 bit offsets 64h + j after the group start, and the walk over
  // codeword starts from offset 0 takes the group's symbols (a group longer
  // than 512 bits takes more steps)
  int s = 0;
  int64_t pos = br.pos;
  if (status == kDecOk) {
    const int eob = nuse + 1;
    const int nsel_kept = min(nsel, kDecSelCap);
    const int cap = c.sym_cap;
    uint16_t* out = sym + c.sym_off;
    int gi = 0, left = kGSize;
    int tbl = uni(sel[0]);
    uint32_t tnext = sel[min(1, nsel_kept - 1)];  // next group's table (read ahead)
    // input words in an LDS ring of 4 chunks of 64 words: chunks cb and
    // cb + 64 are read, cb + 128 is resident and cb + 192 is in flight (pf)
    const int64_t cb0 = pos >> 5;
    int64_t cb = cb0;
    for (int i = lane; i < 192; i += 64) H.ring[i] = gw((uint64_t)(cb0 + i));
    uint32_t pf = gw((uint64_t)(cb0 + 192 + lane));
    __syncthreads();
    for (;;) {
      if (pos > nbits) {
        status = kDecTruncated;
        break;
      }
      if ((pos >> 5) - cb >= 64) {  // at most 80: a step advances at most 532 bits
        H.ring[(cb + 192 - cb0 + lane) & (kRing - 1)] = pf;
        cb += 64;
        pf = gw((uint64_t)(cb + 192 + lane));
        __syncthreads();
      }
      const int sh = (int)(pos & 31);
      const int64_t wrel = (pos >> 5) - cb0;
      const int tb = tbl << kLutBits;
      // phases batched so that the LDS reads of the 8 offsets overlap
      uint32_t wa[8], wb[8], t[8], e[8];
      int nx[8];
#pragma unroll
      for (int h = 0; h < 8; h++) {
        const int64_t q = wrel + ((sh + 64 * h + lane) >> 5);
        wa[h] = H.ring[q & (kRing - 1)];
        wb[h] = H.ring[(q + 1) & (kRing - 1)];
      }
#pragma unroll
      for (int h = 0; h < 8; h++) {
        const int o = sh + 64 * h + lane;
        t[h] = (uint32_t)((((uint64_t)wa[h] << 32) | wb[h]) >> (32 - (o & 31)));
      }
#pragma unroll
      for (int h = 0; h < 8; h++) e[h] = H.lut[tb + (t[h] >> (32 - kLutBits))];
      bool lng = false;
#pragma unroll
      for (int h = 0; h < 8; h++) lng |= (e[h] & 31u) == 0;
      if (ballot(lng)) {
#pragma unroll
        for (int h = 0; h < 8; h++)
          if ((e[h] & 31u) == 0) e[h] = long_code(H, tbl, t[h] >> (32 - kMaxLen));
      }
#pragma unroll
      for (int h = 0; h < 8; h++) {
        // next codeword start; 1023 (ends the walk) after an invalid code or
        // the end-of-block symbol
        const int L = (int)(e[h] & 31u);
        nx[h] = (L == 0 || (int)(e[h] >> 5) == eob) ? 1023 : 64 * h + lane + L;
      }
      // the walk takes whole 64-position segments while the group is not
      // complete (two codewords per step); starts past the group's last
      // symbol are dropped below
      uint64_t M[8];
      int cur = 0, cnt = 0;
#pragma unroll
      for (int h = 0; h < 8; h++) {
        uint64_t m = 0;
        if (cnt < left) {
          const int bnd = 64 * (h + 1);
          while (cur < bnd) {
            m |= 1ull << (cur & 63);
            cur = __builtin_amdgcn_readlane(nx[h], cur & 63);
          }
          cnt += __popcll(m);
        }
        M[h] = m;
      }
      int nw = min(cnt, left);
      bool eob_hit = false, bad = false;
      int64_t npos = pos + cur;
      if (cur == 1023 && cnt - 1 < left) {  // the last start holds an invalid code or EOB
        int last = 0;
#pragma unroll
        for (int h = 0; h < 8; h++)
          if (M[h]) last = 64 * h + 63 - __clzll(M[h]);
        uint32_t el = 0;
#pragma unroll
        for (int h = 0; h < 8; h++)
          if ((last >> 6) == h) el = (uint32_t)__builtin_amdgcn_readlane((int)e[h], last & 63);
        if ((el & 31u) == 0) {
          bad = true;
        } else {
          eob_hit = true;
          nw = cnt - 1;
          npos = pos + last + (int)(el & 31u);
        }
      } else if (cnt > left) {  // the next group starts at the start of rank `left`
        int before = 0, at = -1;
#pragma unroll
        for (int h = 0; h < 8; h++) {
          const uint64_t bl = ballot(((M[h] >> lane) & 1ull) && before + __popcll(M[h] & lanemask_lt()) == left);
          if (at < 0 && bl) at = 64 * h + __ffsll((unsigned long long)bl) - 1;
          before += __popcll(M[h]);
        }
        npos = pos + at;
      }
      if (bad) {
        status = kDecBad;
        break;
      }
      if (s + nw > cap) {
        status = kDecOverflow;
        break;
      }
      int before = 0;
#pragma unroll
      for (int h = 0; h < 8; h++) {
        const int rk = before + __popcll(M[h] & lanemask_lt());
        if (((M[h] >> lane) & 1ull) && rk < nw) out[s + rk] = (uint16_t)(e[h] >> 5);
        before += __popcll(M[h]);
      }
      s += nw;
      pos = npos;
      if (eob_hit) break;
      left -= nw;
      if (left == 0) {
        gi++;
        if (gi >= nsel_kept) {
          status = kDecBad;
          break;
        }
        tbl = uni((int)tnext);
        tnext = sel[min(gi + 1, nsel_kept - 1)];
        left = kGSize;
      }
    }
    if (status == kDecOk && pos > nbits) status = kDecTruncated;
  }
  for (int i = lane; i < 256; i += 64) R->seq[i] = i < nuse ? H.seq[i] : 0;
  if (lane == 0) {
    R->end_bit = pos;
    R->s = s;
    R->status = status;
    R->crc = bcrc;
    R->orig_ptr = orig;
    R->n_in_use = nuse;
    R->next4 = 0;
    R->rest = 0;
  }
}

// ---------------------------------------------------------------- stage 3

// Inverse MTF from the identity list over one tile of symbols, one wave.
// The list lives in the wave's registers, 4 slots per lane (slot of list
// position p in byte p&3 of lane p>>2).  Outputs each symbol's slot, the
// tile's final list and its output length (zero runs expanded).
__global__ void __launch_bounds__(64) k_dec_mtf1(const DecBlk* __restrict__ blk, const int32_t* __restrict__ stblk,
                                                 const uint16_t* __restrict__ sym, uint8_t* __restrict__ slot,
                                                 uint8_t* __restrict__ F, int64_t* __restrict__ tlen,
                                                 DecBlkRes* __restrict__ bres) {
  const int t = blockIdx.x;
  const int b = stblk[t];
  const DecBlk B = blk[b];
  const int i0 = (t - B.st0) * kDecSymTile;
  const int n = min(kDecSymTile, B.s - i0);
  const uint16_t* S = sym + B.sym_off;
  const int lane = threadIdx.x;
  // run symbols (RUNA/RUNB) directly before the tile
  int carry;
  {
    const int j = i0 - 1 - lane;
    const bool isrun = lane < 32 && j >= 0 && S[j] < 2;
    const uint64_t bm = ballot(!isrun);
    carry = __ffsll((unsigned long long)bm) - 1;
  }
  uint32_t wl = (4u * lane) | ((4u * lane + 1u) << 8) | ((4u * lane + 2u) << 16) | ((4u * lane + 3u) << 24);
  int64_t total = 0;
  bool bad = false;
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int i = c0 + lane;
    const bool v = i < n;
    const uint32_t x = v ? S[i0 + i] : 0u;
    const bool nz = v && x >= 2u;
    const uint64_t nzm = ballot(nz), runm = ballot(v && x < 2u);
    if (v) {
      if (nz) {
        total += 1;
      } else {
        const uint64_t before = ~runm & lanemask_lt();
        const int kd = before ? lane - (63 - __clzll(before)) - 1 : carry + lane;
        if (kd >= 24) bad = true;
        else total += (int64_t)(x + 1u) << kd;
      }
    }
    const uint64_t nr = ~runm;
    carry = nr ? 63 - (63 - __clzll(nr)) : carry + 64;
    carry = min(carry, 64);
    const uint32_t front = (uint32_t)__builtin_amdgcn_readfirstlane((int)wl) & 255u;
    uint32_t myslot = 0;
    uint64_t mm = nzm;
    while (mm) {
      const int j = __ffsll((unsigned long long)mm) - 1;
      mm &= mm - 1;
      const uint32_t xi = (uint32_t)__builtin_amdgcn_readlane((int)x, j) - 1u;  // 1..255
      const uint32_t sw = (uint32_t)__builtin_amdgcn_readlane((int)wl, (int)(xi >> 2));
      const uint32_t sl = (sw >> ((xi & 3u) * 8u)) & 255u;
      // list positions 0..xi shift up by one, position 0 takes the slot;
      // the lane's bytes at positions <= xi (the low c of its 4, c clamped
      // to 0..4) take the shifted word: mask = high word of 2^32-1 << 8c
      // (lane 0, which has no lane below, takes sl in the shifted-in byte)
      const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp((int)(sl << 24), (int)wl, 0x138, 0xf, 0xf, false);
      const uint32_t sh = (wl << 8) | (prev >> 24);
      const int c = min(max((int)xi - 4 * lane + 1, 0), 4);
      const uint32_t mask = (uint32_t)((0xFFFFFFFFull << (8 * c)) >> 32);
      wl = (sh & mask) | (wl & ~mask);
      myslot = lane == j ? sl : myslot;
    }
    // a zero run outputs the front slot: that of the last non-run symbol before it
    const uint64_t nzb = nzm & lanemask_lt();
    const int src = nzb ? 63 - __clzll(nzb) : lane;
    const uint32_t got = (uint32_t)__shfl((int)myslot, src, 64);
    const uint32_t so = nz ? myslot : (nzb ? got : front);
    if (v) slot[B.sym_off + i0 + i] = (uint8_t)so;
  }
  reinterpret_cast<uint32_t*>(F + (int64_t)t * 256)[lane] = wl;
  const int64_t tot = wave_sum(total);
  if (lane == 0) tlen[t] = tot;
  if (ballot(bad) && lane == 0) atomicOr(&bres[b].err, 1);
}

// Per block: each tile's starting list (composition of the earlier tiles'
// permutations applied to the block's symbol list), output offsets and m.
__global__ void __launch_bounds__(256) k_dec_mtf_scan(const DecBlk* __restrict__ blk, const DecRes* __restrict__ res,
                                                      const uint8_t* __restrict__ F, uint8_t* __restrict__ P,
                                                      const int64_t* __restrict__ tlen, int64_t* __restrict__ toff,
                                                      DecBlkRes* __restrict__ bres) {
  __shared__ uint8_t cur[256];
  const int b = blockIdx.x;
  const int k = threadIdx.x;
  const DecBlk B = blk[b];
  uint8_t pk = k < B.n_in_use ? res[B.cand].seq[k] : 0;
  int64_t off = 0;
  const int te = B.st0 + B.nst;
  uint8_t f = B.nst ? F[(int64_t)B.st0 * 256 + k] : 0;
  for (int t = B.st0; t < te; t++) {
    P[(int64_t)t * 256 + k] = pk;
    if (k == 0) toff[t] = off;
    off += tlen[t];
    const uint8_t fn = t + 1 < te ? F[(int64_t)(t + 1) * 256 + k] : 0;
    cur[k] = pk;
    __syncthreads();
    pk = cur[f];
    __syncthreads();
    f = fn;
  }
  if (k == 0) {
    int err = 0;
    if (off < 1 || off > kDecMaxM) err |= 2;
    else if (B.orig_ptr >= off) err |= 4;
    bres[b].m = (int32_t)min(off, (int64_t)kDecMaxM + 1);
    if (err) atomicOr(&bres[b].err, err);
  }
}

// Output count of symbol e of a tile whose symbols sit in ss[32 + e] (with
// the 32 symbols before the tile in ss[0..32), 0xFFFF where absent).
__device__ inline int sym_count(const uint16_t* ss, int e) {
  const uint32_t x = ss[32 + e];
  if (x >= 2u) return 1;
  int kd = 0;
  while (kd < 24 && ss[32 + e - 1 - kd] < 2u) kd++;
  return (int)(x + 1u) << kd;
}

// Writes the tile's part of the L column: symbol i covers cum[i]..cum[i+1).
__global__ void __launch_bounds__(256) k_dec_mtf2(const DecBlk* __restrict__ blk, const int32_t* __restrict__ stblk,
                                                  const uint16_t* __restrict__ sym, const uint8_t* __restrict__ slot,
                                                  const uint8_t* __restrict__ P, const int64_t* __restrict__ toff,
                                                  uint8_t* __restrict__ L) {
  __shared__ uint16_t ss[kDecSymTile + 32];
  __shared__ int32_t cum[kDecSymTile + 1];
  __shared__ uint8_t pb[kDecSymTile];
  __shared__ uint8_t pl[256];
  __shared__ int32_t tmp[8];
  const int tid = threadIdx.x;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = stblk[t];
  const DecBlk B = blk[b];
  const int i0 = (t - B.st0) * kDecSymTile;
  const int n = min(kDecSymTile, B.s - i0);
  const uint16_t* S = sym + B.sym_off;
  {
    uint16_t v[17];  // all loads in flight before the LDS writes
#pragma unroll
    for (int q = 0; q < 17; q++) {
      const int e = tid + 256 * q;
      const int j = i0 - 32 + e;
      v[q] = (e < n + 32 && j >= 0) ? S[j] : (uint16_t)0xFFFF;
    }
#pragma unroll
    for (int q = 0; q < 17; q++) {
      const int e = tid + 256 * q;
      if (e < n + 32) ss[e] = v[q];
    }
  }
  pl[tid] = P[(int64_t)t * 256 + tid];
  __syncthreads();
  int c[16];
  int sum = 0;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = tid * 16 + q;
    c[q] = e < n ? sym_count(ss, e) : 0;
    if (e < n) pb[e] = pl[slot[B.sym_off + i0 + e]];
    sum += c[q];
  }
  int tot;
  int run = wg_excl_sum<256, int>(sum, tmp, &tot);
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = tid * 16 + q;
    if (e < n) cum[e] = run;
    run += c[q];
  }
  if (tid == 0) cum[n] = tot;
  __syncthreads();
  uint8_t* dst = L + B.m_off + toff[t];
  for (int o0 = tid * 16; o0 < tot; o0 += 256 * 16) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cum[mid] <= o0) lo = mid;
      else hi = mid - 1;
    }
    int i = lo;
    const int oe = min(o0 + 16, tot);
    for (int o = o0; o < oe; o++) {
      while (cum[i + 1] <= o) i++;
      dst[o] = pb[i];
    }
  }
}

// ---------------------------------------------------------------- stage 4

__global__ void __launch_bounds__(256) k_dec_hist(const DecBlk* __restrict__ blk, const int32_t* __restrict__ mtblk,
                                                  const uint8_t* __restrict__ L, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[4][256];
  const int tid = threadIdx.x;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = mtblk[t];
  const DecBlk B = blk[b];
  const int p0 = (t - B.mt0) * kTile;
  const int n = min(kTile, B.m - p0);
  for (int i = tid; i < 1024; i += 256) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint8_t* src = L + B.m_off + p0;  // 16-byte aligned
  const int w = tid >> 6;
  const int q0 = tid * 16;
  if (q0 + 16 <= n) {
    const uint4 v = *reinterpret_cast<const uint4*>(src + q0);
    const uint32_t x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 16; q++) atomicAdd(&h[w][(x[q >> 2] >> (8 * (q & 3))) & 255u], 1u);
  } else {
    for (int e = q0; e < n; e++) atomicAdd(&h[w][src[e]], 1u);
  }
  __syncthreads();
  hist[(int64_t)t * 256 + tid] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
}

// Per block: hist[t][c] becomes the count of byte c in the block's earlier
// tiles; cb[b][c] = number of bytes < c in the block.
__global__ void __launch_bounds__(256) k_dec_colscan(const DecBlk* __restrict__ blk, uint32_t* __restrict__ hist,
                                                     uint32_t* __restrict__ cb) {
  __shared__ uint32_t tmp[8];
  const int b = blockIdx.x;
  const int c = threadIdx.x;
  const DecBlk B = blk[b];
  uint32_t run = 0;
  for (int t = B.mt0; t < B.mt0 + B.nmt; t++) {
    const uint32_t h = hist[(int64_t)t * 256 + c];
    hist[(int64_t)t * 256 + c] = run;
    run += h;
  }
  uint32_t tot;
  const uint32_t C = wg_excl_sum<256, uint32_t>(run, tmp, &tot);
  cb[b * 256 + c] = C;
}

// Stable placement: L position j with byte c goes to cb[c] + (earlier tiles)
// + (earlier waves of the tile) + (rank within the wave's earlier rounds).
__global__ void __launch_bounds__(256) k_dec_scatter(const DecBlk* __restrict__ blk, const int32_t* __restrict__ mtblk,
                                                     const uint8_t* __restrict__ L, const uint32_t* __restrict__ hist,
                                                     const uint32_t* __restrict__ cb, uint32_t* __restrict__ tt) {
  __shared__ uint8_t lt[kTile];
  __shared__ uint32_t wc[4][256];
  __shared__ uint32_t basec[256];
  const int tid = threadIdx.x;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = mtblk[t];
  const DecBlk B = blk[b];
  const int p0 = (t - B.mt0) * kTile;
  const int n = min(kTile, B.m - p0);
  const uint8_t* src = L + B.m_off + p0;
  for (int e = tid; e < n; e += 256) lt[e] = src[e];
  for (int i = tid; i < 1024; i += 256) (&wc[0][0])[i] = 0;
  basec[tid] = cb[b * 256 + tid] + hist[(int64_t)t * 256 + tid];
  __syncthreads();
  const int w = tid >> 6, lane = tid & 63;
  uint32_t rk[16];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int e = w * 1024 + r * 64 + lane;
    const bool v = e < n;
    const uint32_t c = v ? lt[e] : 0u;
    const uint64_t peers = wave_match_bits<8>(c, ballot(v));
    const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
    const uint32_t before = v ? wc[w][c] : 0u;
    __builtin_amdgcn_wave_barrier();
    if (v && rank == 0) wc[w][c] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rk[r] = before + rank;
  }
  __syncthreads();
  {
    const uint32_t a0 = wc[0][tid], a1 = wc[1][tid], a2 = wc[2][tid];
    wc[0][tid] = 0;
    wc[1][tid] = a0;
    wc[2][tid] = a0 + a1;
    wc[3][tid] = a0 + a1 + a2;
  }
  __syncthreads();
  uint32_t* dst = tt + B.m_off;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int e = w * 1024 + r * 64 + lane;
    if (e < n) {
      const uint32_t c = lt[e];
      dst[basec[c] + wc[w][c] + rk[r]] = ((uint32_t)(p0 + e) << 8) | c;
    }
  }
}

// ---------------------------------------------------------------- stage 5

__device__ inline int origin_node(const DecBlk& B) {
  const int nsp = (B.m + kDecSplit - 1) / kDecSplit;
  return B.orig_ptr % kDecSplit == 0 ? B.orig_ptr / kDecSplit : nsp;
}
__device__ inline uint32_t node_row(const DecBlk& B, int i) {
  const int nsp = (B.m + kDecSplit - 1) / kDecSplit;
  return i < nsp ? (uint32_t)(i * kDecSplit) : (uint32_t)B.orig_ptr;
}

__global__ void __launch_bounds__(256) k_dec_walk1(const DecBlk* __restrict__ blk, const int32_t* __restrict__ nodeblk,
                                                   int32_t nnodes, const uint32_t* __restrict__ tt,
                                                   int32_t* __restrict__ nx, int32_t* __restrict__ ln,
                                                   uint8_t* __restrict__ stg, uint32_t* __restrict__ rrow,
                                                   DecBlkRes* __restrict__ bres) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= nnodes) return;
  const int b = nodeblk[g];
  const DecBlk B = blk[b];
  const uint32_t* T = tt + B.m_off;
  const uint32_t orig = (uint32_t)B.orig_ptr;
  uint64_t* S = reinterpret_cast<uint64_t*>(stg + (int64_t)g * kDecStage);
  uint32_t r = node_row(B, g - B.node0);
  int len = 0;
  uint64_t acc = 0;
  for (;;) {
    if (len == kDecStage) rrow[g] = r;
    const uint32_t e = T[r];
    if (len < kDecStage) {
      acc |= (uint64_t)(e & 255u) << (8 * (len & 7));
      if ((len & 7) == 7) {
        S[len >> 3] = acc;
        acc = 0;
      }
    }
    r = e >> 8;
    len++;
    if (r % kDecSplit == 0 || r == orig || len > B.m) break;
  }
  const int kept = min(len, kDecStage);
  if (kept & 7) S[kept >> 3] = acc;
  if (len > B.m) atomicOr(&bres[b].err, 16);
  nx[g] = r == orig ? origin_node(B) : (int32_t)(r / kDecSplit);
  ln[g] = len;
}

// Per block: splitter positions by pointer jumping on the list that starts
// at origPtr's node (its predecessor is cut off).
__global__ void __launch_bounds__(512) k_dec_rank(const DecBlk* __restrict__ blk, const int32_t* __restrict__ nx,
                                                  const int32_t* __restrict__ ln, int32_t* __restrict__ pos,
                                                  DecBlkRes* __restrict__ bres) {
  __shared__ int32_t snx[kDecMaxNodes];
  __shared__ int32_t sd[kDecMaxNodes];
  constexpr int PER = (kDecMaxNodes + 511) / 512;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const DecBlk B = blk[b];
  const int nn = B.nnode;
  const int org = origin_node(B);
  for (int i = tid; i < nn; i += 512) {
    const int v = nx[B.node0 + i];
    snx[i] = v == org ? -1 : v;
    sd[i] = ln[B.node0 + i];
  }
  __syncthreads();
  for (int round = 0; round < 13; round++) {
    int nn_[PER], nd_[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const int i = tid + j * 512;
      nn_[j] = -1;
      nd_[j] = 0;
      if (i < nn) {
        const int a = snx[i];
        nn_[j] = a < 0 ? -1 : snx[a];
        nd_[j] = a < 0 ? sd[i] : sd[i] + sd[a];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const int i = tid + j * 512;
      if (i < nn) {
        snx[i] = nn_[j];
        sd[i] = nd_[j];
      }
    }
    __syncthreads();
  }
  // nodes that never reach the cut lie on other cycles (periodic block)
  const int cyc = snx[org] < 0 ? sd[org] : 0;
  for (int i = tid; i < nn; i += 512) pos[B.node0 + i] = snx[i] < 0 ? cyc - sd[i] : -1;
  if (tid == 0) {
    bres[b].cyc = cyc;
    if (cyc < 1 || cyc > B.m) atomicOr(&bres[b].err, 32);
  }
}

// Staged bytes to their place: 16 bytes per thread, kDecStage / 16 threads
// per walker.
__global__ void __launch_bounds__(256) k_dec_copy(const DecBlk* __restrict__ blk, const int32_t* __restrict__ nodeblk,
                                                  int32_t nnodes, const uint8_t* __restrict__ stg,
                                                  const int32_t* __restrict__ pos, const int32_t* __restrict__ ln,
                                                  uint8_t* __restrict__ T, const DecBlkRes* __restrict__ bres) {
  constexpr int kChunks = kDecStage / 16;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int g = (int)(t / kChunks), ch = (int)(t % kChunks);
  if (g >= nnodes) return;
  const int p = pos[g], len = ln[g];
  const int n = min(len, kDecStage);
  if (p < 0 || ch * 16 >= n) return;
  const int b = nodeblk[g];
  if (bres[b].err || p + len > bres[b].cyc) return;
  const uint4 v = reinterpret_cast<const uint4*>(stg + (int64_t)g * kDecStage)[ch];
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
  uint8_t* dst = T + blk[b].m_off + p + ch * 16;
  const int c = min(16, n - ch * 16);
#pragma unroll
  for (int i = 0; i < 16; i++)
    if (i < c) dst[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
}

// Stretches longer than kDecStage: the bytes past the staged ones, walked
// again from the row the first walk reached there.
__global__ void __launch_bounds__(256) k_dec_walk2(const DecBlk* __restrict__ blk, const int32_t* __restrict__ nodeblk,
                                                   int32_t nnodes, const uint32_t* __restrict__ tt,
                                                   const int32_t* __restrict__ pos, const int32_t* __restrict__ ln,
                                                   const uint32_t* __restrict__ rrow, uint8_t* __restrict__ T,
                                                   const DecBlkRes* __restrict__ bres) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= nnodes) return;
  const int len = ln[g];
  if (len <= kDecStage) return;
  const int b = nodeblk[g];
  if (bres[b].err) return;
  const DecBlk B = blk[b];
  const uint32_t* M = tt + B.m_off;
  const int p = pos[g];
  if (p < 0 || p + len > bres[b].cyc) return;
  uint8_t* dst = T + B.m_off + p;
  uint32_t r = rrow[g];
  for (int k = kDecStage; k < len; k++) {
    const uint32_t e = M[r];
    dst[k] = (uint8_t)e;
    r = e >> 8;
  }
}

// Periodic blocks: T[p] = T[p mod cyc] beyond the first cycle.
__global__ void __launch_bounds__(256) k_dec_repl(const DecBlk* __restrict__ blk, const int32_t* __restrict__ mtblk,
                                                  uint8_t* __restrict__ T, const DecBlkRes* __restrict__ bres) {
  const int t = blockIdx.x;
  const int b = mtblk[t];
  const int cyc = bres[b].cyc;
  const DecBlk B = blk[b];
  if (bres[b].err || cyc >= B.m || cyc < 1) return;
  const int p0 = (t - B.mt0) * kTile;
  const int n = min(kTile, B.m - p0);
  uint8_t* D = T + B.m_off;
  for (int e = threadIdx.x; e < n; e += 256) {
    const int p = p0 + e;
    if (p >= cyc) D[p] = D[p % cyc];
  }
}

// ---------------------------------------------------------------- stage 6
// RLE1 state machine; a transfer function packs, per entry state, the exit
// state (bits 24..26) and the bytes output (bits 0..23).

// f[s] for a state s in 0..4, as masks (a select chain here is turned into
// an indexed stack array by the compiler)
__device__ inline uint32_t pick5(const uint32_t f[5], uint32_t s) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 5; k++) r |= f[k] & (0u - (uint32_t)(s == (uint32_t)k));
  return r;
}

// r = a then b
__device__ inline void compose5(const uint32_t a[5], const uint32_t b[5], uint32_t r[5]) {
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const uint32_t x = a[s];
    const uint32_t y = pick5(b, x >> 24);
    r[s] = (y & 0xFF000000u) | ((x & 0xFFFFFFu) + (y & 0xFFFFFFu));
  }
}

__device__ inline void rle1_fn(const uint8_t x[16], int cnt, uint32_t pb, uint32_t f[5]) {
#pragma unroll
  for (int s0 = 0; s0 < 5; s0++) {
    uint32_t s = (uint32_t)s0, p = pb, out = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (q < cnt) {
        const uint32_t v = x[q];
        if (s == 4u) {
          out += v;
          s = 0;
        } else if (s > 0u && v == p) {
          s++;
          out++;
        } else {
          s = 1;
          out++;
        }
        p = v;
      }
    }
    f[s0] = (s << 24) | out;
  }
}

// Loads the thread's 16 bytes of a T tile and the byte before them.
__device__ inline int load_t16(const uint8_t* T, int p0, int n, int tid, uint8_t x[16], uint32_t* pb) {
  const int q0 = tid * 16;
  const int c = max(0, min(16, n - q0));
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = k < c ? T[p0 + q0 + k] : 0;
  *pb = (p0 + q0 > 0) ? T[p0 + q0 - 1] : 0u;
  return c;
}

// Inclusive wave scan of transfer functions (lane order).
__device__ inline void wave_fn_scan(uint32_t f[5]) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t g[5], r[5];
#pragma unroll
    for (int s = 0; s < 5; s++) g[s] = (uint32_t)__shfl_up((int)f[s], d, 64);
    compose5(g, f, r);
    if (lane >= d) {
#pragma unroll
      for (int s = 0; s < 5; s++) f[s] = r[s];
    }
  }
}

__global__ void __launch_bounds__(256) k_dec_rle1a(const DecBlk* __restrict__ blk, const int32_t* __restrict__ mtblk,
                                                   const uint8_t* __restrict__ T, uint32_t* __restrict__ tf) {
  __shared__ uint32_t wf[4][5];
  const int tid = threadIdx.x;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = mtblk[t];
  const DecBlk B = blk[b];
  const int p0 = (t - B.mt0) * kTile;
  const int n = min(kTile, B.m - p0);
  uint8_t x[16];
  uint32_t pb;
  const int c = load_t16(T + B.m_off, p0, n, tid, x, &pb);
  uint32_t f[5];
  rle1_fn(x, c, pb, f);
  wave_fn_scan(f);
  if (lane_id() == 63) {
#pragma unroll
    for (int s = 0; s < 5; s++) wf[tid >> 6][s] = f[s];
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t acc[5], r[5];
#pragma unroll
    for (int s = 0; s < 5; s++) acc[s] = wf[0][s];
    for (int w = 1; w < 4; w++) {
      uint32_t g[5];
#pragma unroll
      for (int s = 0; s < 5; s++) g[s] = wf[w][s];
      compose5(acc, g, r);
#pragma unroll
      for (int s = 0; s < 5; s++) acc[s] = r[s];
    }
#pragma unroll
    for (int s = 0; s < 5; s++) tf[(int64_t)t * 5 + s] = acc[s];
  }
}

// Per block (one wave): entry state and output offset of every tile, and
// the block's raw length.
__global__ void __launch_bounds__(64) k_dec_rle1scan(const DecBlk* __restrict__ blk, const uint32_t* __restrict__ tf,
                                                     int32_t* __restrict__ tent, int64_t* __restrict__ too,
                                                     DecBlkRes* __restrict__ bres) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const DecBlk B = blk[b];
  uint32_t s = 0;
  int64_t off = 0;
  for (int c0 = 0; c0 < B.nmt; c0 += 64) {
    const int t = B.mt0 + c0 + lane;
    const bool v = c0 + lane < B.nmt;
    uint32_t f[5];
#pragma unroll
    for (int k = 0; k < 5; k++) f[k] = v ? tf[(int64_t)t * 5 + k] : 0u;
    const int cnt = min(64, B.nmt - c0);
    int32_t my_s = 0;
    int64_t my_off = 0;
    for (int j = 0; j < cnt; j++) {
      const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)pick5(f, s), j);
      if (lane == j) {
        my_s = (int32_t)s;
        my_off = off;
      }
      off += e & 0xFFFFFFu;
      s = e >> 24;
    }
    if (v) {
      tent[t] = my_s;
      too[t] = my_off;
    }
  }
  if (lane == 0) bres[b].n = off;
}

// Writes the raw bytes of one T tile and its CRC register (from zero, over
// the tile's output).
__global__ void __launch_bounds__(256) k_dec_rle1b(const DecBlk* __restrict__ blk, const int32_t* __restrict__ mtblk,
                                                   const uint8_t* __restrict__ T, const int32_t* __restrict__ tent,
                                                   const int64_t* __restrict__ too, uint8_t* __restrict__ out,
                                                   uint32_t* __restrict__ treg) {
  __shared__ uint32_t wf[4][5];
  __shared__ int32_t cum[kTile + 1];
  __shared__ uint8_t eb[kTile];
  __shared__ uint32_t tab[4][256];
  __shared__ int32_t tmp[8];
  __shared__ uint32_t red[4];
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = mtblk[t];
  const DecBlk B = blk[b];
  const int p0 = (t - B.mt0) * kTile;
  const int n = min(kTile, B.m - p0);
#pragma unroll
  for (int k = 0; k < 4; k++) tab[k][tid] = c_crc4.s[k][tid];
  uint8_t x[16];
  uint32_t pb;
  const int c = load_t16(T + B.m_off, p0, n, tid, x, &pb);
  uint32_t f[5];
  rle1_fn(x, c, pb, f);
  wave_fn_scan(f);
  if (lane == 63) {
#pragma unroll
    for (int s = 0; s < 5; s++) wf[w][s] = f[s];
  }
  __syncthreads();
  // entry state of this thread: tile entry, then earlier waves, then earlier lanes
  uint32_t s = (uint32_t)tent[t];
  for (int k = 0; k < w; k++) s = pick5(wf[k], s) >> 24;
  {
    uint32_t g[5];
#pragma unroll
    for (int k = 0; k < 5; k++) g[k] = (uint32_t)__shfl_up((int)f[k], 1, 64);
    if (lane > 0) s = pick5(g, s) >> 24;
  }
  // replay with the known state: bytes output per position and which byte
  int cnt[16];
  int sum = 0;
  {
    uint32_t p = pb;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      cnt[q] = 0;
      if (q < c) {
        const uint32_t v = x[q];
        uint32_t e = v;
        if (s == 4u) {
          cnt[q] = (int)v;
          e = p;
          s = 0;
        } else if (s > 0u && v == p) {
          s++;
          cnt[q] = 1;
        } else {
          s = 1;
          cnt[q] = 1;
        }
        eb[tid * 16 + q] = (uint8_t)e;
        p = v;
      }
      sum += cnt[q];
    }
  }
  int tot;
  int run = wg_excl_sum<256, int>(sum, tmp, &tot);
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = tid * 16 + q;
    if (e < n) cum[e] = run;
    run += cnt[q];
  }
  if (tid == 0) cum[n] = tot;
  __syncthreads();
  uint8_t* dst = out + B.raw_off + too[t];
  uint32_t acc = 0;
  for (int o0 = tid * 16; o0 < tot; o0 += 256 * 16) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cum[mid] <= o0) lo = mid;
      else hi = mid - 1;
    }
    int i = lo;
    const int oe = min(o0 + 16, tot);
    uint32_t reg = 0;
    if (oe - o0 == 16) {
      // the 16 bytes in registers, then the register four bytes per step
      uint32_t wd[4] = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int k = 0; k < 16; k++) {
        while (cum[i + 1] <= o0 + k) i++;
        const uint32_t v = eb[i];
        dst[o0 + k] = (uint8_t)v;
        wd[k >> 2] |= v << (24 - 8 * (k & 3));
      }
#pragma unroll
      for (int w4 = 0; w4 < 4; w4++) {
        const uint32_t v = reg ^ wd[w4];
        reg = tab[3][v >> 24] ^ tab[2][(v >> 16) & 255u] ^ tab[1][(v >> 8) & 255u] ^ tab[0][v & 255u];
      }
    } else {
      for (int o = o0; o < oe; o++) {
        while (cum[i + 1] <= o) i++;
        const uint32_t v = eb[i];
        dst[o] = (uint8_t)v;
        reg = (reg << 8) ^ tab[0][(reg >> 24) ^ v];
      }
    }
    acc ^= crc_shift(reg, (uint32_t)(tot - oe));
  }
  for (int d = 32; d >= 1; d >>= 1) acc ^= (uint32_t)__shfl_xor((int)acc, d, 64);
  if (lane == 0) red[w] = acc;
  __syncthreads();
  if (tid == 0) treg[t] = red[0] ^ red[1] ^ red[2] ^ red[3];
}

// Per block (one wave): the block CRC from the tile registers.
__global__ void __launch_bounds__(64) k_dec_crc(const DecBlk* __restrict__ blk, const uint32_t* __restrict__ treg,
                                                const int64_t* __restrict__ too, DecBlkRes* __restrict__ bres) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const DecBlk B = blk[b];
  const int64_t n = B.n;
  uint32_t acc = 0;
  for (int k = lane; k < B.nmt; k += 64) {
    const int t = B.mt0 + k;
    const int64_t end = k + 1 < B.nmt ? too[t + 1] : n;
    acc ^= crc_shift(treg[t], (uint32_t)(n - end));
  }
  for (int d = 32; d >= 1; d >>= 1) acc ^= (uint32_t)__shfl_xor((int)acc, d, 64);
  if (lane == 0) {
    const uint32_t crc = crc_final(acc, (uint32_t)n);
    bres[b].crc = crc;
    if (crc != B.crc_want) atomicOr(&bres[b].err, 64);
  }
}

inline unsigned grid_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ======================================================================= host

Decoder::Decoder(int device) : device_(device) {
  (void)hipSetDevice(device_);
  (void)hipHostMalloc((void**)&h_small_, 64 * sizeof(int32_t), hipHostMallocDefault);
  for (auto& e : ev_) (void)hipEventCreate(&e);
}

Decoder::~Decoder() {
  (void)hipSetDevice(device_);
  Buf* bufs[] = {&w_, &cand_, &res_, &sym_, &slot_, &idx_, &small_, &blk_, &bres_, &stblk_, &f_, &p_, &tlen_, &toff_,
                 &l_, &tt_, &mtblk_, &hist_, &cb_, &nodeblk_, &nx_, &ln_, &pos_, &t_, &tf_, &tent_, &too_, &treg_, &out_,
                 &stg_, &rrow_};
  for (Buf* b : bufs)
    if (b->p) (void)hipFree(b->p);
  if (h_small_) (void)hipHostFree(h_small_);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
}

int Decoder::need(Buf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (bytes <= b.cap) return 0;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    if (e != hipSuccess) return (int)e;
  }
  const size_t want = bytes + bytes / 8 + 256;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) return (int)e;
  b.cap = want;
  return 0;
}

int Decoder::fail(const std::string& msg) {
  err_ = msg;
  return kDecDataError;
}

// launch with the per-kernel event probe (benchmarking aid) around it
#define DEC_LAUNCH_AS(name, k, grid, block, shm, s, ...)  \
  do {                                                     \
    probe_begin(name, s);                                  \
    hipLaunchKernelGGL(k, grid, block, shm, s, __VA_ARGS__); \
    probe_end(name, s);                                    \
  } while (0)
#define DEC_LAUNCH(k, grid, block, shm, s, ...) DEC_LAUNCH_AS(#k, k, grid, block, shm, s, __VA_ARGS__)
#define DEC_CK(x)                         \
  do {                                    \
    hipError_t e_ = (x);                  \
    if (e_ != hipSuccess) return (int)e_; \
  } while (0)
#define DEC_NEED(b, n)          \
  do {                          \
    int r_ = need((b), (n));    \
    if (r_) return r_;          \
  } while (0)

int Decoder::decode_cands(const std::vector<int32_t>& idx, hipStream_t s) {
  if (idx.empty()) return 0;
  DEC_NEED(idx_, idx.size() * sizeof(int32_t));
  DEC_CK(hipMemcpyAsync(idx_.p, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  // 12-bit tables while every block's decode fits on the CUs at two per
  // CU, 9-bit tables (eleven per CU) for more blocks
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_) != hipSuccess || cus < 1) cus = 256;
  if (idx.size() <= (size_t)(2 * cus)) {
    DEC_LAUNCH_AS("k_dec_huff", k_dec_huff<12>, dim3((unsigned)idx.size()), dim3(64), 0, s, (const uint32_t*)w_.p, nbits_,
               (nbits_ + 31) / 32 + kPadWords, (const DecCand*)cand_.p, (const int32_t*)idx_.p, (DecRes*)res_.p,
               (uint16_t*)sym_.p, (uint8_t*)sel_.p);
  } else {
    DEC_LAUNCH_AS("k_dec_huff", k_dec_huff<9>, dim3((unsigned)idx.size()), dim3(64), 0, s, (const uint32_t*)w_.p, nbits_,
               (nbits_ + 31) / 32 + kPadWords, (const DecCand*)cand_.p, (const int32_t*)idx_.p, (DecRes*)res_.p,
               (uint16_t*)sym_.p, (uint8_t*)sel_.p);
  }
  return (int)hipGetLastError();
}

static const char* status_text(int st) {
  switch (st) {
    case kDecOverflow: return "symbol overflow";
    case kDecBad: return "malformed block header or code";
    case kDecRandomised: return "randomised blocks are not supported";
    case kDecTruncated: return "truncated block";
    default: return "unknown block error";
  }
}

int Decoder::run(const uint8_t* d_in, size_t len, hipStream_t s, DecodeTimes* tm) {
  err_.clear();
  out_len_ = 0;
  DEC_CK(hipSetDevice(device_));
  if (tm) *tm = DecodeTimes{};
  if (len == 0) return 0;
  uint8_t hdr[4] = {0, 0, 0, 0};
  DEC_CK(hipMemcpyAsync(hdr, d_in, std::min<size_t>(4, len), hipMemcpyDeviceToHost, s));
  DEC_CK(hipStreamSynchronize(s));
  auto is_hdr = [](const uint8_t* h) { return h[0] == 'B' && h[1] == 'Z' && h[2] == 'h' && h[3] >= '1' && h[3] <= '9'; };
  if (len < 4 || !is_hdr(hdr)) return fail("not a bzip2 stream (bad stream header)");
  nbits_ = (int64_t)len * 8;
  const int64_t nwords = ((int64_t)len + 3) / 4;
  DEC_CK(hipEventRecord(ev_[0], s));
  // stage 1: words and magic candidates
  DEC_NEED(w_, (size_t)(nwords + kPadWords) * 4);
  DEC_LAUNCH(k_dec_words, dim3(grid_for(nwords + kPadWords, 256)), dim3(256), 0, s, d_in, (int64_t)len,
                     (uint32_t*)w_.p, nwords + kPadWords);
  const int32_t cap = (int32_t)std::min<int64_t>((int64_t)len / 16 + 1024, 1 << 28);
  DEC_NEED(cand_, (size_t)cap * sizeof(DecCand));
  DEC_NEED(small_, 256);
  DEC_CK(hipMemsetAsync(small_.p, 0, 256, s));
  DEC_LAUNCH(k_dec_magic, dim3(grid_for(nwords, 256)), dim3(256), 0, s, (const uint32_t*)w_.p, nbits_,
                     (DecCand*)cand_.p, cap, (int32_t*)small_.p);
  DEC_CK(hipGetLastError());
  DEC_CK(hipMemcpyAsync(h_small_, small_.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DEC_CK(hipStreamSynchronize(s));
  const int32_t ncand = h_small_[0];
  if (ncand > cap) return fail("too many block-magic candidates");
  cands_.resize(ncand);
  if (ncand) DEC_CK(hipMemcpy(cands_.data(), cand_.p, (size_t)ncand * sizeof(DecCand), hipMemcpyDeviceToHost));
  std::sort(cands_.begin(), cands_.end(), [](const DecCand& a, const DecCand& b) { return a.bit < b.bit; });
  // symbol slots: a block's symbols are at most its bits, and at most kDecMaxM
  int64_t slab_max = kDecMaxM;
  if (const char* e = getenv("BZG_DEC_SLAB_MAX")) slab_max = std::max<int64_t>(1, std::min<int64_t>(kDecMaxM, atoll(e)));
  auto assign = [&](const std::vector<char>* big) {
    sym_tot_ = 0;
    for (int32_t i = 0; i < ncand; i++) {
      DecCand& c = cands_[i];
      c.sym_off = sym_tot_;
      c.sym_cap = 0;
      if (c.kind != 0) continue;
      const int64_t gap = (i + 1 < ncand ? cands_[i + 1].bit : nbits_) - c.bit;
      int64_t k = std::min<int64_t>(slab_max, gap);
      if (big && (*big)[i]) k = kDecMaxM;
      c.sym_cap = (int32_t)k;
      sym_tot_ += (k + 7) & ~7ll;
    }
  };
  assign(nullptr);
  std::vector<int32_t> all(ncand);
  for (int32_t i = 0; i < ncand; i++) all[i] = i;
  res_h_.resize(ncand);
  bool retried = false;
  // stage 2 (and once more with full slots for candidates that overflowed)
  for (;;) {
    DEC_NEED(sym_, (size_t)sym_tot_ * 2);
    DEC_NEED(slot_, (size_t)sym_tot_);
    DEC_NEED(res_, (size_t)std::max(ncand, 1) * sizeof(DecRes));
    DEC_NEED(sel_, (size_t)std::max(ncand, 1) * 2 * kDecSelCap);
    if (ncand) DEC_CK(hipMemcpyAsync(cand_.p, cands_.data(), (size_t)ncand * sizeof(DecCand), hipMemcpyHostToDevice, s));
    if (int r = decode_cands(all, s)) return r;
    if (ncand) DEC_CK(hipMemcpyAsync(res_h_.data(), res_.p, (size_t)ncand * sizeof(DecRes), hipMemcpyDeviceToHost, s));
    DEC_CK(hipStreamSynchronize(s));
    bool over = false;
    std::vector<char> big(ncand, 0);
    for (int32_t i = 0; i < ncand; i++)
      if (cands_[i].kind == 0 && res_h_[i].status == kDecOverflow) big[i] = over = true;
    if (!over || retried) break;
    retried = true;
    assign(&big);
  }
  DEC_CK(hipEventRecord(ev_[1], s));
  // the chain of blocks: stream header, blocks back to back, end-of-stream
  std::vector<DecBlk> blks;
  int streams = 0;
  {
    auto find = [&](int64_t bit) -> int32_t {
      auto it = std::lower_bound(cands_.begin(), cands_.end(), bit,
                                 [](const DecCand& c, int64_t v) { return c.bit < v; });
      return (it != cands_.end() && it->bit == bit) ? (int32_t)(it - cands_.begin()) : -1;
    };
    int64_t byte = 0;
    for (;;) {
      streams++;
      int64_t bit = (byte + 4) * 8;
      uint32_t combined = 0;
      for (;;) {
        const int32_t ci = find(bit);
        if (ci < 0) {
          char m[96];
          snprintf(m, sizeof m, "no block or end-of-stream magic at bit %lld", (long long)bit);
          return fail(m);
        }
        const DecRes& r = res_h_[ci];
        if (r.status != kDecOk) {
          char m[128];
          snprintf(m, sizeof m, "block at bit %lld: %s", (long long)bit, status_text(r.status));
          return fail(m);
        }
        if (cands_[ci].kind == 1) {
          if (r.crc != combined) return fail("stream CRC mismatch");
          byte = (r.end_bit + 7) >> 3;
          break;
        }
        DecBlk B{};
        B.sym_off = cands_[ci].sym_off;
        B.s = r.s;
        B.orig_ptr = r.orig_ptr;
        B.n_in_use = r.n_in_use;
        B.cand = ci;
        B.crc_want = r.crc;
        blks.push_back(B);
        combined = ((combined << 1) | (combined >> 31)) ^ r.crc;
        bit = r.end_bit;
      }
      const DecRes& eos = res_h_[find(bit)];
      if (eos.rest <= 0) break;
      uint8_t nx4[4] = {(uint8_t)(eos.next4 >> 24), (uint8_t)(eos.next4 >> 16), (uint8_t)(eos.next4 >> 8),
                        (uint8_t)eos.next4};
      if (eos.rest >= 4 && is_hdr(nx4)) continue;
      static const uint8_t kHead[3] = {'B', 'Z', 'h'};
      if (eos.rest < 4 && memcmp(nx4, kHead, (size_t)std::min<int64_t>(eos.rest, 3)) == 0)
        return fail("truncated stream header after end of stream");
      break;  // trailing bytes that are no stream are ignored
    }
  }
  const int32_t nb = (int32_t)blks.size();
  if (tm) {
    tm->blocks = nb;
    tm->streams = streams;
  }
  if (nb == 0) {
    DEC_CK(hipStreamSynchronize(s));
    return 0;
  }
  // stage 3: inverse MTF
  std::vector<int32_t> stblk;
  for (int32_t j = 0; j < nb; j++) {
    blks[j].st0 = (int32_t)stblk.size();
    blks[j].nst = (blks[j].s + kDecSymTile - 1) / kDecSymTile;
    for (int k = 0; k < blks[j].nst; k++) stblk.push_back(j);
  }
  const int32_t nst = (int32_t)stblk.size();
  DEC_NEED(blk_, (size_t)nb * sizeof(DecBlk));
  DEC_NEED(bres_, (size_t)nb * sizeof(DecBlkRes));
  DEC_NEED(stblk_, (size_t)std::max(nst, 1) * 4);
  DEC_NEED(f_, (size_t)std::max(nst, 1) * 256);
  DEC_NEED(p_, (size_t)std::max(nst, 1) * 256);
  DEC_NEED(tlen_, (size_t)std::max(nst, 1) * 8);
  DEC_NEED(toff_, (size_t)std::max(nst, 1) * 8);
  DEC_CK(hipMemcpyAsync(blk_.p, blks.data(), (size_t)nb * sizeof(DecBlk), hipMemcpyHostToDevice, s));
  if (nst) DEC_CK(hipMemcpyAsync(stblk_.p, stblk.data(), (size_t)nst * 4, hipMemcpyHostToDevice, s));
  DEC_CK(hipMemsetAsync(bres_.p, 0, (size_t)nb * sizeof(DecBlkRes), s));
  if (nst)
    DEC_LAUNCH(k_dec_mtf1, dim3(nst), dim3(64), 0, s, (const DecBlk*)blk_.p, (const int32_t*)stblk_.p,
                       (const uint16_t*)sym_.p, (uint8_t*)slot_.p, (uint8_t*)f_.p, (int64_t*)tlen_.p,
                       (DecBlkRes*)bres_.p);
  DEC_LAUNCH(k_dec_mtf_scan, dim3(nb), dim3(256), 0, s, (const DecBlk*)blk_.p, (const DecRes*)res_.p,
                     (const uint8_t*)f_.p, (uint8_t*)p_.p, (const int64_t*)tlen_.p, (int64_t*)toff_.p,
                     (DecBlkRes*)bres_.p);
  DEC_CK(hipGetLastError());
  std::vector<DecBlkRes> br(nb);
  DEC_CK(hipMemcpyAsync(br.data(), bres_.p, (size_t)nb * sizeof(DecBlkRes), hipMemcpyDeviceToHost, s));
  DEC_CK(hipStreamSynchronize(s));
  for (int32_t j = 0; j < nb; j++)
    if (br[j].err) {
      char m[96];
      snprintf(m, sizeof m, "block %d: bad run length, block length or origPtr (code %d)", j, br[j].err);
      return fail(m);
    }
  // layouts over m
  std::vector<int32_t> mtblk, nodeblk;
  int64_t m_tot = 0;
  for (int32_t j = 0; j < nb; j++) {
    DecBlk& B = blks[j];
    B.m = br[j].m;
    B.m_off = m_tot;
    m_tot += ((int64_t)B.m + 15) & ~15ll;
    B.mt0 = (int32_t)mtblk.size();
    B.nmt = (B.m + kTile - 1) / kTile;
    for (int k = 0; k < B.nmt; k++) mtblk.push_back(j);
    B.node0 = (int32_t)nodeblk.size();
    B.nnode = (B.m + kDecSplit - 1) / kDecSplit + (B.orig_ptr % kDecSplit != 0 ? 1 : 0);
    for (int k = 0; k < B.nnode; k++) nodeblk.push_back(j);
  }
  const int32_t nmt = (int32_t)mtblk.size(), nnodes = (int32_t)nodeblk.size();
  DEC_NEED(l_, (size_t)m_tot + 64);
  DEC_NEED(t_, (size_t)m_tot + 64);
  DEC_NEED(tt_, (size_t)m_tot * 4 + 64);
  DEC_NEED(mtblk_, (size_t)nmt * 4);
  DEC_NEED(hist_, (size_t)nmt * 256 * 4);
  DEC_NEED(cb_, (size_t)nb * 256 * 4);
  DEC_NEED(nodeblk_, (size_t)nnodes * 4);
  DEC_NEED(nx_, (size_t)nnodes * 4);
  DEC_NEED(ln_, (size_t)nnodes * 4);
  DEC_NEED(pos_, (size_t)nnodes * 4);
  DEC_NEED(stg_, (size_t)nnodes * kDecStage);
  DEC_NEED(rrow_, (size_t)nnodes * 4);
  DEC_NEED(tf_, (size_t)nmt * 5 * 4);
  DEC_NEED(tent_, (size_t)nmt * 4);
  DEC_NEED(too_, (size_t)nmt * 8);
  DEC_NEED(treg_, (size_t)nmt * 4);
  DEC_CK(hipMemcpyAsync(blk_.p, blks.data(), (size_t)nb * sizeof(DecBlk), hipMemcpyHostToDevice, s));
  DEC_CK(hipMemcpyAsync(mtblk_.p, mtblk.data(), (size_t)nmt * 4, hipMemcpyHostToDevice, s));
  DEC_CK(hipMemcpyAsync(nodeblk_.p, nodeblk.data(), (size_t)nnodes * 4, hipMemcpyHostToDevice, s));
  const DecBlk* dB = (const DecBlk*)blk_.p;
  DEC_LAUNCH(k_dec_mtf2, dim3(nst), dim3(256), 0, s, dB, (const int32_t*)stblk_.p, (const uint16_t*)sym_.p,
                     (const uint8_t*)slot_.p, (const uint8_t*)p_.p, (const int64_t*)toff_.p, (uint8_t*)l_.p);
  DEC_CK(hipEventRecord(ev_[2], s));
  // stage 4: FL map
  DEC_LAUNCH(k_dec_hist, dim3(nmt), dim3(256), 0, s, dB, (const int32_t*)mtblk_.p, (const uint8_t*)l_.p,
                     (uint32_t*)hist_.p);
  DEC_LAUNCH(k_dec_colscan, dim3(nb), dim3(256), 0, s, dB, (uint32_t*)hist_.p, (uint32_t*)cb_.p);
  DEC_LAUNCH(k_dec_scatter, dim3(nmt), dim3(256), 0, s, dB, (const int32_t*)mtblk_.p, (const uint8_t*)l_.p,
                     (const uint32_t*)hist_.p, (const uint32_t*)cb_.p, (uint32_t*)tt_.p);
  DEC_CK(hipEventRecord(ev_[3], s));
  // stage 5: inverse BWT
  DEC_LAUNCH(k_dec_walk1, dim3(grid_for(nnodes, 256)), dim3(256), 0, s, dB, (const int32_t*)nodeblk_.p, nnodes,
                     (const uint32_t*)tt_.p, (int32_t*)nx_.p, (int32_t*)ln_.p, (uint8_t*)stg_.p, (uint32_t*)rrow_.p,
                     (DecBlkRes*)bres_.p);
  DEC_LAUNCH(k_dec_rank, dim3(nb), dim3(512), 0, s, dB, (const int32_t*)nx_.p, (const int32_t*)ln_.p,
                     (int32_t*)pos_.p, (DecBlkRes*)bres_.p);
  DEC_LAUNCH(k_dec_copy, dim3(grid_for((int64_t)nnodes * (kDecStage / 16), 256)), dim3(256), 0, s, dB,
             (const int32_t*)nodeblk_.p, nnodes, (const uint8_t*)stg_.p, (const int32_t*)pos_.p,
             (const int32_t*)ln_.p, (uint8_t*)t_.p, (const DecBlkRes*)bres_.p);
  DEC_LAUNCH(k_dec_walk2, dim3(grid_for(nnodes, 256)), dim3(256), 0, s, dB, (const int32_t*)nodeblk_.p, nnodes,
             (const uint32_t*)tt_.p, (const int32_t*)pos_.p, (const int32_t*)ln_.p, (const uint32_t*)rrow_.p,
             (uint8_t*)t_.p, (const DecBlkRes*)bres_.p);
  DEC_LAUNCH(k_dec_repl, dim3(nmt), dim3(256), 0, s, dB, (const int32_t*)mtblk_.p, (uint8_t*)t_.p,
                     (const DecBlkRes*)bres_.p);
  DEC_CK(hipEventRecord(ev_[4], s));
  // stage 6: inverse RLE1
  DEC_LAUNCH(k_dec_rle1a, dim3(nmt), dim3(256), 0, s, dB, (const int32_t*)mtblk_.p, (const uint8_t*)t_.p,
                     (uint32_t*)tf_.p);
  DEC_LAUNCH(k_dec_rle1scan, dim3(nb), dim3(64), 0, s, dB, (const uint32_t*)tf_.p, (int32_t*)tent_.p,
                     (int64_t*)too_.p, (DecBlkRes*)bres_.p);
  DEC_CK(hipGetLastError());
  DEC_CK(hipMemcpyAsync(br.data(), bres_.p, (size_t)nb * sizeof(DecBlkRes), hipMemcpyDeviceToHost, s));
  DEC_CK(hipStreamSynchronize(s));
  int64_t raw = 0;
  for (int32_t j = 0; j < nb; j++) {
    if (br[j].err) {
      char m[96];
      snprintf(m, sizeof m, "block %d: inverse BWT failed (code %d)", j, br[j].err);
      return fail(m);
    }
    if (br[j].n > 0x7FFFFFFFll) return fail("block expands beyond 2 GiB");
    blks[j].raw_off = raw;
    blks[j].n = (int32_t)br[j].n;
    raw += br[j].n;
  }
  DEC_NEED(out_, (size_t)raw + 64);
  DEC_CK(hipMemcpyAsync(blk_.p, blks.data(), (size_t)nb * sizeof(DecBlk), hipMemcpyHostToDevice, s));
  DEC_LAUNCH(k_dec_rle1b, dim3(nmt), dim3(256), 0, s, dB, (const int32_t*)mtblk_.p, (const uint8_t*)t_.p,
                     (const int32_t*)tent_.p, (const int64_t*)too_.p, (uint8_t*)out_.p, (uint32_t*)treg_.p);
  DEC_LAUNCH(k_dec_crc, dim3(nb), dim3(64), 0, s, dB, (const uint32_t*)treg_.p, (const int64_t*)too_.p,
                     (DecBlkRes*)bres_.p);
  DEC_CK(hipGetLastError());
  DEC_CK(hipEventRecord(ev_[5], s));
  DEC_CK(hipMemcpyAsync(br.data(), bres_.p, (size_t)nb * sizeof(DecBlkRes), hipMemcpyDeviceToHost, s));
  DEC_CK(hipStreamSynchronize(s));
  for (int32_t j = 0; j < nb; j++)
    if (br[j].err) {
      char m[96];
      snprintf(m, sizeof m, "block %d: CRC mismatch (stored %08x, computed %08x)", j, blks[j].crc_want, br[j].crc);
      return fail(m);
    }
  out_len_ = (size_t)raw;
  if (tm) {
    float a = 0;
    (void)hipEventElapsedTime(&a, ev_[0], ev_[1]);
    tm->huff = a;
    (void)hipEventElapsedTime(&a, ev_[1], ev_[2]);
    tm->mtf = a;
    (void)hipEventElapsedTime(&a, ev_[2], ev_[3]);
    tm->sort = a;
    (void)hipEventElapsedTime(&a, ev_[3], ev_[4]);
    tm->walk = a;
    (void)hipEventElapsedTime(&a, ev_[4], ev_[5]);
    tm->rle1 = a;
  }
  return 0;
}

}  // namespace bzg
