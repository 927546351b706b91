// front.hip - per-block state reset, block CRC and RLE1 for a whole batch.
//
// CRC-32/BZIP2 of each raw block (reference: src/compression/crc32.cpp:22-29,
// called per block at src/compression/compress.cu:259) is computed tile-
// parallel: every thread forms the CRC register of its 16 bytes started from
// zero and shifts it by the bytes that follow it
// (reg(A||B) = reg(A) * x^(8|B|) + reg0(B) mod P), so a tile's register is
// the XOR of its threads' and a block's the XOR of its shifted tiles'.
//
// RLE1 (reference: src/compression/rle1.cu:13-113) is restated per byte
// position: a run [a, e) is cut in pieces of 259; offset r = (i-a) mod 259
// emits the byte for r < 3, the byte and the piece's count byte for r == 3,
// nothing beyond.  The number of bytes a position emits depends only on its
// distance to the run start, so one pass over the raw tiles gives each tile's
// output count (positions before the tile's first run start are settled by
// the per-block scan from the carried run start), and the write pass places
// bytes in input order.
#include "crc.h"
#include "pipeline.h"

namespace bzg {

// Bytes emitted by the first j positions of a run: 5 per 259-piece, and
// min(o, 3) or 5 for the o positions of a started piece.
__device__ inline int rle1_prefix(int j) {
  int q = j / 259, o = j - 259 * q;
  return 5 * q + (o <= 3 ? o : 5);
}

__global__ void __launch_bounds__(256) k_block_init(Dev d) {
  int b = blockIdx.x;
  BlockState* st = d.st + b;
  if (threadIdx.x == 0) {
    st->crc = 0; st->m = 0; st->s = 0; st->orig_ptr = 0; st->n_in_use = 0;
    st->n_groups = 0; st->n_sel = 0; st->rounds = 0; st->bits = 0;
    st->hdr_bits = 0; st->bit_off = 0; st->err = 0; st->rk = 0;
  }
  if (threadIdx.x < 8) st->present[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < 260; i += blockDim.x) d.freq[(int64_t)b * 260 + i] = 0;
  if (b == 0 && threadIdx.x < CTR_N) d.ctr[threadIdx.x] = 0;
}

// ------------------------------------------------- wave-per-tile layout
// The CRC / RLE1 tile passes run one wave per raw tile, four tiles per
// 256-thread workgroup; lane l holds bytes [64 l, 64 l + 64) of its tile as
// 16 little-endian words and works on them four bytes at a time.  Byte
// flags are words with 0x80 in each flagged byte.
constexpr int kLaneBytes = kTile / 64;
static_assert(kLaneBytes == 64, "a lane holds 16 words of its tile");
constexpr uint32_t kHi = 0x80808080u;

// Mask of the low r bytes of a word (r clamped to 0..4).
__device__ inline uint32_t byte_mask(int r) { return r >= 4 ? ~0u : (r <= 0 ? 0u : (1u << (8 * r)) - 1u); }

// Bytes [p, p + c), 0 <= c <= 64, as 16 words, zero past c.  Only the
// 16-byte aligned chunks that hold bytes of the range are read, so no read
// leaves the pages of the range.  The range's offset within its chunk is
// the same in every lane of the wave (lanes are 64 bytes apart), so the
// word shift is a uniform branch.
__device__ inline void load64(const uint8_t* p, int c, uint32_t w[16]) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  const uintptr_t a0 = a & ~(uintptr_t)15;
  uint32_t u[20];
#pragma unroll
  for (int j = 0; j < 5; j++) {
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (c > 0 && a0 + 16 * j < a + (uintptr_t)c) v = *reinterpret_cast<const uint4*>(a0 + 16 * j);
    u[4 * j] = v.x;
    u[4 * j + 1] = v.y;
    u[4 * j + 2] = v.z;
    u[4 * j + 3] = v.w;
  }
  const int sh = __builtin_amdgcn_readfirstlane((int)(a & 15));
  // bytes sh .. sh+63 of u: shift by sh mod 4 bytes, then by 1 and 2 words
  // (selects with constant indexes keep u in registers)
  uint32_t v[19];
#pragma unroll
  for (int j = 0; j < 19; j++) v[j] = __builtin_amdgcn_alignbyte(u[j + 1], u[j], (uint32_t)(sh & 3));
  // masked blends (the empty asm hides that the masks are uniform, which
  // would otherwise turn the shifts into a scratch copy at a scalar offset)
  uint32_t m1 = (sh & 4) ? ~0u : 0u, m2 = (sh & 8) ? ~0u : 0u;
  asm volatile("" : "+v"(m1), "+v"(m2));
#pragma unroll
  for (int j = 0; j < 18; j++) v[j] ^= (v[j] ^ v[j + 1]) & m1;
#pragma unroll
  for (int i = 0; i < 16; i++) w[i] = v[i] ^ ((v[i] ^ v[i + 2]) & m2);
  if (c < 64) {
#pragma unroll
    for (int i = 0; i < 16; i++) w[i] &= byte_mask(c - 4 * i);
  }
}

// 0x80 in every nonzero byte of x.
__device__ inline uint32_t nz80(uint32_t x) { return (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & kHi; }

// Byte flags of the lane's first r bytes of a word (r clamped to 0..4).
__device__ inline uint32_t low_flags(int r) { return r >= 4 ? kHi : (r <= 0 ? 0u : kHi & ((1u << (8 * r)) - 1u)); }

// Run starts of a lane's bytes: byte j starts a run when it differs from the
// byte before it (prev: the byte before the lane's first), or when
// first_is_start (the block's first byte); bytes at or past c never do.
__device__ inline void run_flags(const uint32_t w[16], uint32_t prev, bool first_is_start, int c, uint32_t D[16]) {
  uint32_t lo = prev << 24;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    D[i] = nz80(w[i] ^ __builtin_amdgcn_alignbyte(w[i], lo, 3u));
    lo = w[i];
  }
  if (first_is_start) D[0] |= 0x80u;
  if (c < 64) {
#pragma unroll
    for (int i = 0; i < 16; i++) D[i] &= low_flags(c - 4 * i);
  }
}

// First and last flagged byte of a lane (64 / -1 when none).
__device__ inline void flag_bounds(const uint32_t D[16], int& fi, int& li) {
  fi = 64;
  li = -1;
#pragma unroll
  for (int i = 15; i >= 0; i--)
    if (D[i]) fi = 4 * i + (__builtin_ctz(D[i]) >> 3);
#pragma unroll
  for (int i = 0; i < 16; i++)
    if (D[i]) li = 4 * i + ((31 - __builtin_clz(D[i])) >> 3);
}

// A run's output restarts every 259 bytes (a piece of 4 bytes and a count).
// For a lane whose first byte lies d0 >= 1 bytes after the start of the run
// it continues, the piece starts of that run act as run starts: one may
// fall inside the lane (added to E before the lane's first real start fi)
// and one may lie in the 4 bytes before the lane (returned as the flags of
// those bytes, byte 3 being the byte just before the lane).
__device__ inline uint32_t carried_starts(int d0, int fi, uint32_t E[16]) {
  const int dm = d0 % 259;  // offset of the lane's first byte within its piece
  int jv = dm == 0 ? 0 : 259 - dm;
  if (jv >= fi) jv = 64;  // only the carried run's piece starts
  // a flag per word from its own offset (the empty asm keeps the compiler
  // from turning the word loop into one store at a computed index)
  asm volatile("" : "+v"(jv));
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t r = (uint32_t)(jv - 4 * i);
    E[i] |= r < 4u ? 0x80u << (8 * r) : 0u;
  }
  return dm >= 1 && dm <= 4 ? 0x80u << (8 * (4 - dm)) : 0u;
}

// Output bytes of four positions with run-start flags E (bytes of word i)
// and Em (the word before): a position whose offset in its piece is below 3
// emits its byte (lt3), offset 3 emits its byte and the count (eq3), larger
// offsets nothing.
__device__ inline void piece_flags(uint32_t E, uint32_t Em, uint32_t& lt3, uint32_t& eq3) {
  const uint32_t s1 = __builtin_amdgcn_alignbyte(E, Em, 3u);
  const uint32_t s2 = __builtin_amdgcn_alignbyte(E, Em, 2u);
  const uint32_t s3 = __builtin_amdgcn_alignbyte(E, Em, 1u);
  const uint32_t o = E | s1 | s2;
  lt3 = o & kHi;
  eq3 = s3 & ~o & kHi;
}

// Tile pass 1 (one wave per tile): CRC register of the tile (from zero),
// first/last run start (block-relative) and the RLE1 output count of the
// positions whose run starts inside the tile.
__global__ void __launch_bounds__(256) k_front1(Dev d) {
  __shared__ uint32_t tab[4][256];
  const int tid = threadIdx.x, lane = lane_id();
#pragma unroll
  for (int k = 0; k < 4; k++) tab[k][tid] = c_crc4.s[k][tid];
  __syncthreads();
  const int t = 4 * (int)xcd_remap(blockIdx.x, gridDim.x) + (tid >> 6);
  if (t >= d.ntile_n) return;
  const int b = d.tn_blk[t];
  const BlockDesc D = d.desc[b];
  const int k = t - D.tile_n0;
  const int start = k * kTile;
  const int len = min(kTile, D.n - start);
  const uint8_t* src = d.in + D.in_off + start;
  const int q0 = lane * kLaneBytes;
  const int c = max(0, min(kLaneBytes, len - q0));
  uint32_t w[16];
  load64(src + q0, c, w);
  // CRC register of the lane's bytes from zero, moved to the tile's end
  uint32_t reg = 0;
  if (c == kLaneBytes) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t v = reg ^ __builtin_bswap32(w[i]);
      reg = tab[3][v >> 24] ^ tab[2][(v >> 16) & 255u] ^ tab[1][(v >> 8) & 255u] ^ tab[0][v & 255u];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 16; i++) {
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (4 * i + q < c) reg = (reg << 8) ^ tab[0][(reg >> 24) ^ ((w[i] >> (8 * q)) & 255u)];
    }
  }
  if (c > 0) reg = crc_shift(reg, (uint32_t)(len - q0 - c));
  // run starts
  uint32_t prev = (uint32_t)__shfl_up((int)(w[15] >> 24), 1, 64);
  if (lane == 0) prev = start > 0 ? src[-1] : 0u;
  uint32_t E[16];
  run_flags(w, prev, lane == 0 && start == 0, c, E);
  int fi, li;
  flag_bounds(E, fi, li);
  // last run start before the lane inside the tile (tile-relative)
  const int lbl = li >= 0 ? q0 + li : -1;
  int rp = __shfl_up(wave_incl_max(lbl), 1, 64);
  if (lane == 0) rp = -1;
  // positions whose run starts inside the tile: all of the lane's when a
  // start precedes it in the tile, else those from its first start on
  uint32_t em = 0;
  if (rp >= 0 && c > 0) em = carried_starts(q0 - rp, fi, E);
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    uint32_t lt3, eq3;
    piece_flags(E[i], em, lt3, eq3);
    em = E[i];
    const uint32_t cm = low_flags(c - 4 * i) & (rp >= 0 ? kHi : ~low_flags(fi - 4 * i));
    cnt += __popc(lt3 & cm) + 2 * __popc(eq3 & cm);
  }
  reg = wave_xor(reg);
  cnt = wave_sum(cnt);
  const int fb = wave_min(fi < 64 ? start + q0 + fi : 0x7fffffff);
  const int lb = wave_max(li >= 0 ? start + q0 + li : -1);
  if (lane == 0) {
    d.tn_crc[t] = reg;
    d.tn_cnt[t] = cnt;
    d.tn_fb[t] = fb;
    d.tn_lb[t] = lb;
  }
}

// Per-block pass over the tile summaries: block CRC, RLE1 length m, each
// tile's output offset, and the carried run bounds the write pass needs
// (tn_lb <- last run start before the tile, tn_fb <- first run start at or
// after the tile end, or n).
__global__ void __launch_bounds__(256) k_front_scan(Dev d) {
  __shared__ int32_t tmp[8];
  __shared__ int32_t xs[256];
  __shared__ uint32_t red[4];
  const int tid = threadIdx.x, lane = lane_id();
  const int b = blockIdx.x;
  const BlockDesc D = d.desc[b];
  const int nt = tiles_for(D.n);
  int run_lb = -1, run_cnt = 0;
  uint32_t acc = 0;
  for (int c0 = 0; c0 < nt; c0 += 256) {
    const int tk = c0 + tid;
    const bool v = tk < nt;
    const int ti = D.tile_n0 + tk;
    const int lb = v ? d.tn_lb[ti] : -1;
    const int fb = v ? d.tn_fb[ti] : 0x7fffffff;
    int cnt = v ? d.tn_cnt[ti] : 0;
    const uint32_t reg = v ? d.tn_crc[ti] : 0u;
    xs[tid] = wg_incl_max<256, int>(lb, tmp, run_lb);
    __syncthreads();
    const int ca = tid > 0 ? xs[tid - 1] : run_lb;
    const int next_lb = xs[255];
    const int start = tk * kTile, end = min(start + kTile, D.n);
    if (v && ca >= 0) {
      const int pe = min(fb, end);  // positions [start, pe) continue the run from ca
      if (pe > start) cnt += rle1_prefix(pe - ca) - rle1_prefix(start - ca);
    }
    int tot;
    const int ex = wg_excl_sum<256, int>(cnt, tmp, &tot);
    if (v) {
      d.tn_off[ti] = run_cnt + ex;
      d.tn_lb[ti] = ca;
      acc ^= crc_shift(reg, (uint32_t)(D.n - end));
    }
    run_cnt += tot;
    run_lb = next_lb;
    __syncthreads();
  }
  int run_fb = D.n;
  for (int c0 = ((nt - 1) / 256) * 256; c0 >= 0 && nt > 0; c0 -= 256) {
    const int tk = c0 + tid;
    const bool v = tk < nt;
    const int ti = D.tile_n0 + tk;
    const int fb = v ? d.tn_fb[ti] : 0x7fffffff;
    xs[tid] = wg_rev_incl_min<256, int>(fb, tmp, run_fb);
    __syncthreads();
    const int ce = tid < 255 ? xs[tid + 1] : run_fb;
    const int next_fb = xs[0];
    if (v) d.tn_fb[ti] = ce;
    run_fb = next_fb;
    __syncthreads();
  }
  for (int o = 32; o >= 1; o >>= 1) acc ^= __shfl_xor(acc, o, 64);
  if (lane == 0) red[tid >> 6] = acc;
  __syncthreads();
  if (tid == 0) {
    const uint32_t reg = red[0] ^ red[1] ^ red[2] ^ red[3];
    d.st[b].m = run_cnt;
    d.st[b].crc = ~(gf2_mulmod(0xFFFFFFFFu, xpow8((uint32_t)D.n)) ^ reg);
  }
}

// RLE1 write pass (one wave per tile): the tile's output bytes are staged
// in LDS in input order, then copied to the block's RLE1 text with word
// stores (byte stores at the two edge words shared with the neighbouring
// tiles).  A word whose four positions each emit their byte is staged
// whole; only words with a count byte or a silent position go byte by byte.
constexpr int kWaveOut = kTile + kTile / 4 + 16;  // a tile's output (<= 5/4) and slack
static_assert(kWaveOut % 16 == 0, "staging rows stay word-aligned");
__global__ void __launch_bounds__(256) k_rle1_write(Dev d) {
  __shared__ __align__(16) uint8_t obuf[4][kWaveOut];
  const int tid = threadIdx.x, lane = lane_id();
  const int t = 4 * (int)xcd_remap(blockIdx.x, gridDim.x) + (tid >> 6);
  if (t >= d.ntile_n) return;
  uint8_t* ob = obuf[tid >> 6];
  const int b = d.tn_blk[t];
  const BlockDesc D = d.desc[b];
  const int k = t - D.tile_n0;
  const int start = k * kTile;
  const int len = min(kTile, D.n - start);
  const uint8_t* src = d.in + D.in_off + start;
  // last run start before the tile, first run start at or after its end
  const int ca = d.tn_lb[t], ce = d.tn_fb[t];
  const int q0 = lane * kLaneBytes;
  const int c = max(0, min(kLaneBytes, len - q0));
  const int p0 = start + q0;  // block-relative position of the lane's first byte
  uint32_t w[16];
  load64(src + q0, c, w);
  uint32_t prev = (uint32_t)__shfl_up((int)(w[15] >> 24), 1, 64);
  if (lane == 0) prev = start > 0 ? src[-1] : 0u;
  uint32_t E[16];
  run_flags(w, prev, lane == 0 && start == 0, c, E);
  int fi, li;
  flag_bounds(E, fi, li);
  // next real run start after each word (lane-relative, may lie beyond the
  // lane), for the count bytes
  int rn = __shfl_down(wave_incl_min_rev(fi < 64 ? p0 + fi : 0x7fffffff), 1, 64);
  if (lane == 63) rn = 0x7fffffff;
  rn = min(rn, ce);
  int nx[16];
  {
    int nxt = rn - p0;
#pragma unroll
    for (int i = 15; i >= 0; i--) {
      nx[i] = nxt;
      if (E[i]) nxt = 4 * i + (__builtin_ctz(E[i]) >> 3);
    }
  }
  // last run start before the lane (block-relative)
  int rp = __shfl_up(wave_incl_max(li >= 0 ? p0 + li : -1), 1, 64);
  if (lane == 0) rp = -1;
  rp = max(rp, ca);
  uint32_t em0 = 0;
  uint32_t R[16];  // real run starts (E gains the carried run's piece starts)
#pragma unroll
  for (int i = 0; i < 16; i++) R[i] = E[i];
  if (rp >= 0 && c > 0) em0 = carried_starts(p0 - rp, fi, E);
  int cnt = 0;
  {
    uint32_t em = em0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      uint32_t lt3, eq3;
      piece_flags(E[i], em, lt3, eq3);
      em = E[i];
      const uint32_t vm = low_flags(c - 4 * i);
      cnt += __popc(lt3 & vm) + 2 * __popc(eq3 & vm);
    }
  }
  const int incl = wave_incl_sum(cnt);
  int o = incl - cnt;
  const int total = __builtin_amdgcn_readlane(incl, 63);
  {
    uint32_t em = em0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      uint32_t lt3, eq3;
      piece_flags(E[i], em, lt3, eq3);
      em = E[i];
      const uint32_t vm = low_flags(c - 4 * i);
      const uint32_t L = lt3 & vm, Q = eq3 & vm;
      if (L == kHi) {
        if ((o & 3) == 0) {
          *reinterpret_cast<uint32_t*>(ob + o) = w[i];
        } else {
#pragma unroll
          for (int q = 0; q < 4; q++) ob[o + q] = (uint8_t)(w[i] >> (8 * q));
        }
        o += 4;
      } else if (L | Q) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t bit = 0x80u << (8 * q);
          if ((L | Q) & bit) ob[o++] = (uint8_t)(w[i] >> (8 * q));
          if (Q & bit) {
            // count of the piece: bytes after its first four up to the run's
            // end, at most 255 (src/compression/rle1.cu:13-58)
            const uint32_t after = q < 3 ? R[i] >> (8 * q + 8) : 0u;
            const int end = after ? 4 * i + q + 1 + (__builtin_ctz(after) >> 3) : nx[i];
            ob[o++] = (uint8_t)min(255, end - (4 * i + q) - 1);
          }
        }
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint8_t* dst = d.T + D.off + d.tn_off[t];
  const uintptr_t A0 = reinterpret_cast<uintptr_t>(dst) & ~(uintptr_t)3;
  const int lead = (int)(reinterpret_cast<uintptr_t>(dst) - A0);
  const int nw = (lead + total + 3) >> 2;
  for (int wi = lane; wi < nw; wi += 64) {
    const int oo = 4 * wi - lead;
    if (oo >= 0 && oo + 4 <= total) {
      // the 4 bytes from two aligned words (the staging row has slack)
      const uint32_t* o32 = reinterpret_cast<const uint32_t*>(ob) + (oo >> 2);
      const uint32_t v = __builtin_amdgcn_alignbyte(o32[1], o32[0], (uint32_t)(oo & 3));
      *reinterpret_cast<uint32_t*>(A0 + 4 * (uintptr_t)wi) = v;
    } else {
      for (int q = 0; q < 4; q++) {
        const int oq = oo + q;
        if (oq >= 0 && oq < total) dst[oq] = ob[oq];
      }
    }
  }
}

// Per-block exclusive scan of the RLE2 counts of the capacity tiles -> st.s
// (adds the end-of-block symbol).
__global__ void __launch_bounds__(256) k_tile_scan_c(Dev d) {
  __shared__ int32_t tmp[8];
  const int b = blockIdx.x;
  const BlockDesc D = d.desc[b];
  const int t0 = D.tile_c0;
  const int nt = tiles_for(D.cap);
  int run = 0;
  for (int base = 0; base < nt; base += 256) {
    int j = base + threadIdx.x;
    int v = j < nt ? d.tc_cnt[t0 + j] : 0;
    int tot;
    int ex = wg_excl_sum<256, int>(v, tmp, &tot);
    if (j < nt) d.tc_off[t0 + j] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) {
    // end-of-block symbol = n_in_use + 1 (src/compression/rle2.cu:71-74)
    int eob = d.st[b].n_in_use + 1;
    d.SYM[D.off + run] = (uint16_t)eob;
    atomicAdd(&d.freq[(int64_t)b * 260 + eob], 1);
    d.st[b].s = run + 1;
  }
}

void launch_block_init(const Dev& d, hipStream_t s) {
  k_block_init<<<d.nblk, 256, 0, s>>>(d);
}

void launch_crc_rle1(const Dev& d, hipStream_t s) {
  const int wgs = (d.ntile_n + 3) / 4;  // four tiles per workgroup, one per wave
  if (d.ntile_n > 0) k_front1<<<wgs, 256, 0, s>>>(d);
  k_front_scan<<<d.nblk, 256, 0, s>>>(d);
  if (d.ntile_n > 0) k_rle1_write<<<wgs, 256, 0, s>>>(d);
}

void launch_tile_scan_c(const Dev& d, hipStream_t s) {
  k_tile_scan_c<<<d.nblk, 256, 0, s>>>(d);
}

}  // namespace bzg
