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
    st->hdr_bits = 0; st->bit_off = 0; st->err = 0; st->pad_ = 0;
  }
  if (threadIdx.x < 8) st->present[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < 260; i += blockDim.x) d.freq[(int64_t)b * 260 + i] = 0;
  if (b == 0 && threadIdx.x < CTR_N) d.ctr[threadIdx.x] = 0;
}

// Loads the up-to-16 bytes of thread `tid` in a tile; returns the count.
__device__ inline int load16(const uint8_t* src, int len, int tid, uint8_t x[16]) {
  int q0 = tid * 16;
  int c = len - q0;
  c = c < 0 ? 0 : (c > 16 ? 16 : c);
  if (c == 16 && ((reinterpret_cast<uintptr_t>(src + q0) & 15) == 0)) {
    uint4 v = *reinterpret_cast<const uint4*>(src + q0);
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = (uint8_t)(w[k >> 2] >> (8 * (k & 3)));
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = k < c ? src[q0 + k] : 0;
  }
  return c;
}

// Tile pass 1: CRC register of the tile (from zero), first/last run start
// (block-relative) and the RLE1 output count of the positions whose run
// starts inside the tile.
__global__ void __launch_bounds__(256) k_front1(Dev d) {
  __shared__ uint32_t tab[4][256];
  __shared__ uint8_t lastb[256];
  __shared__ int32_t xs[256];
  __shared__ int32_t tmp[8];
  __shared__ uint32_t red[16];
  const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tn_blk[t];
  const BlockDesc D = d.desc[b];
  const int k = t - D.tile_n0;
  const int start = k * kTile;
  const int len = min(kTile, D.n - start);
  const uint8_t* src = d.in + D.in_off + start;
#pragma unroll
  for (int k = 0; k < 4; k++) tab[k][tid] = c_crc4.s[k][tid];
  uint8_t x[16];
  const int c = load16(src, len, tid, x);
  lastb[tid] = c > 0 ? x[c - 1] : 0;
  __syncthreads();
  // (fixed-count loops over x[] keep the register indexing static)
  uint32_t reg = 0;
  if (c == 16) {
    // four bytes per step, the four table reads of a step independent
#pragma unroll
    for (int w = 0; w < 4; w++) {
      const uint32_t v = reg ^ (((uint32_t)x[4 * w] << 24) | ((uint32_t)x[4 * w + 1] << 16) |
                                ((uint32_t)x[4 * w + 2] << 8) | (uint32_t)x[4 * w + 3]);
      reg = tab[3][v >> 24] ^ tab[2][(v >> 16) & 255u] ^ tab[1][(v >> 8) & 255u] ^ tab[0][v & 255u];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 16; q++)
      if (q < c) reg = (reg << 8) ^ tab[0][(reg >> 24) ^ x[q]];
  }
  if (c > 0) reg = crc_shift(reg, (uint32_t)(len - tid * 16 - c));
  // run starts
  const int p0 = start + tid * 16;
  uint32_t bmask = 0;
  {
    uint32_t prev = 0;
    bool has_prev = false;
    if (tid > 0) { prev = lastb[tid - 1]; has_prev = true; }
    else if (start > 0) { prev = src[-1]; has_prev = true; }
    if (c > 0 && (!has_prev || x[0] != prev)) bmask = 1u;
#pragma unroll
    for (int q = 1; q < 16; q++)
      if (q < c && x[q] != x[q - 1]) bmask |= 1u << q;
  }
  int fb = bmask ? p0 + __builtin_ctz(bmask) : 0x7fffffff;
  int lb = bmask ? p0 + 31 - __builtin_clz(bmask) : -1;
  xs[tid] = wg_incl_max<256, int>(lb, tmp, -1);
  __syncthreads();
  // offset r = (i - run start) mod 259, carried from position to position;
  // -1 while the run start lies before the tile
  const int run = tid > 0 ? xs[tid - 1] : -1;
  int r = run >= 0 ? (p0 - run) % 259 : -1;
  int cnt = 0;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    if (q > 0 && r >= 0) r = r == 258 ? 0 : r + 1;
    if (bmask & (1u << q)) r = 0;
    if (q < c && r >= 0) cnt += r < 3 ? 1 : (r == 3 ? 2 : 0);
  }
  // reductions: xor of registers, sum of counts, min/max run starts
  for (int o = 32; o >= 1; o >>= 1) {
    reg ^= __shfl_xor(reg, o, 64);
    cnt += __shfl_xor(cnt, o, 64);
  }
  fb = wave_min(fb);
  lb = wave_max(lb);
  if (lane == 0) {
    red[wv] = reg;
    red[4 + wv] = (uint32_t)cnt;
    red[8 + wv] = (uint32_t)fb;
    red[12 + wv] = (uint32_t)lb;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t r2 = red[0];
    int c2 = (int)red[4], f = (int)red[8], l = (int)red[12];
    for (int w = 1; w < 4; w++) {
      r2 ^= red[w];
      c2 += (int)red[4 + w];
      f = min(f, (int)red[8 + w]);
      l = max(l, (int)red[12 + w]);
    }
    d.tn_crc[t] = r2;
    d.tn_cnt[t] = c2;
    d.tn_fb[t] = f;
    d.tn_lb[t] = l;
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

// RLE1 write pass: every position's bytes are staged in LDS in input order,
// then copied to the block's RLE1 text with word stores (byte stores at the
// two edge words shared with the neighbouring tiles).  A tile where every
// position emits exactly its own byte (no run reaches 4) is staged as it was
// loaded.
__global__ void __launch_bounds__(256) k_rle1_write(Dev d) {
  __shared__ uint8_t lastb[256];
  __shared__ int32_t tmp[8];
  __shared__ int32_t xa[256], xe[256];
  __shared__ __align__(16) uint8_t ob[2 * kTile + 16];
  const int tid = threadIdx.x;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tn_blk[t];
  const BlockDesc D = d.desc[b];
  const int k = t - D.tile_n0;
  const int start = k * kTile;
  const int len = min(kTile, D.n - start);
  const uint8_t* src = d.in + D.in_off + start;
  const int ca = d.tn_lb[t], ce = d.tn_fb[t];
  uint8_t x[16];
  const int c = load16(src, len, tid, x);
  lastb[tid] = c > 0 ? x[c - 1] : 0;
  __syncthreads();
  const int p0 = start + tid * 16;
  uint32_t prev = 0;
  bool has_prev = false;
  if (tid > 0) { prev = lastb[tid - 1]; has_prev = true; }
  else if (start > 0) { prev = src[-1]; has_prev = true; }
  uint32_t bmask = 0;
  if (c > 0 && (!has_prev || x[0] != prev)) bmask = 1u;
#pragma unroll
  for (int q = 1; q < 16; q++)
    if (q < c && x[q] != x[q - 1]) bmask |= 1u << q;
  const int lastB = bmask ? p0 + 31 - __builtin_clz(bmask) : -1;
  const int firstB = bmask ? p0 + __builtin_ctz(bmask) : 0x7fffffff;
  xa[tid] = wg_incl_max<256, int>(lastB, tmp, ca);
  xe[tid] = wg_rev_incl_min<256, int>(firstB, tmp, ce);
  __syncthreads();
  int a = tid > 0 ? xa[tid - 1] : ca;
  int cnt = 0;
  bool ident = true;  // every position emits its own byte, nothing else
  int8_t rr[16];
  {
    // offset in the 259-piece, carried from position to position
    int r = c > 0 && a >= 0 ? (p0 - a) % 259 : 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (q > 0) r = r == 258 ? 0 : r + 1;
      if (bmask & (1u << q)) r = 0;
      const int rq = q < c ? r : 4;
      rr[q] = (int8_t)(rq < 4 ? rq : 4);
      cnt += rq < 3 ? 1 : (rq == 3 ? 2 : 0);
      if (q < c && rq >= 3) ident = false;
    }
  }
  int total;
  int o = wg_excl_sum<256, int>(cnt, tmp, &total);
  if (__syncthreads_and(ident)) {
    // identity tile: thread i's bytes are output bytes 16i..16i+c
    uint32_t w[4];
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++)
      w[k2] = (uint32_t)x[4 * k2] | (uint32_t)x[4 * k2 + 1] << 8 | (uint32_t)x[4 * k2 + 2] << 16 |
              (uint32_t)x[4 * k2 + 3] << 24;
    *reinterpret_cast<uint4*>(ob + 16 * tid) = make_uint4(w[0], w[1], w[2], w[3]);
  } else {
    // run end of every position: next run start after it
    int e = tid < 255 ? xe[tid + 1] : ce;
    uint32_t ends[16];
#pragma unroll
    for (int q = 15; q >= 0; q--) {
      ends[q] = (uint32_t)e;
      if (bmask & (1u << q)) e = p0 + q;
    }
    a = tid > 0 ? xa[tid - 1] : ca;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (bmask & (1u << q)) a = p0 + q;
      if (rr[q] < 4) {
        ob[o++] = x[q];
        if (rr[q] == 3) {
          // j = 259*piece + 3: a full piece when 259*(piece+1) <= run length,
          // else the last piece, whose count is run length - (j - 3) - 4
          const int runlen = (int)ends[q] - a, j = p0 + q - a;
          ob[o++] = j + 256 <= runlen ? 255 : (uint8_t)(runlen - j - 1);
        }
      }
    }
  }
  __syncthreads();
  uint8_t* dst = d.T + D.off + d.tn_off[t];
  const uintptr_t A0 = reinterpret_cast<uintptr_t>(dst) & ~(uintptr_t)3;
  const int lead = (int)(reinterpret_cast<uintptr_t>(dst) - A0);
  const int nw = (lead + total + 3) >> 2;
  for (int w = tid; w < nw; w += 256) {
    const int oo = 4 * w - lead;
    if (oo >= 0 && oo + 4 <= total) {
      // the 4 bytes from two aligned words (ob has 4 bytes of slack)
      const uint32_t* o32 = reinterpret_cast<const uint32_t*>(ob) + (oo >> 2);
      const uint32_t v = (uint32_t)((((uint64_t)o32[1] << 32) | o32[0]) >> (8 * (oo & 3)));
      *reinterpret_cast<uint32_t*>(A0 + 4 * (uintptr_t)w) = v;
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
  if (d.ntile_n > 0) k_front1<<<d.ntile_n, 256, 0, s>>>(d);
  k_front_scan<<<d.nblk, 256, 0, s>>>(d);
  if (d.ntile_n > 0) k_rle1_write<<<d.ntile_n, 256, 0, s>>>(d);
}

void launch_tile_scan_c(const Dev& d, hipStream_t s) {
  k_tile_scan_c<<<d.nblk, 256, 0, s>>>(d);
}

}  // namespace bzg
