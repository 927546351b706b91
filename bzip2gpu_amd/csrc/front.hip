// front.hip - per-block state reset, block CRC and RLE1 for a whole batch.
//
// CRC-32/BZIP2 of each raw block (reference: src/compression/crc32.cpp:22-29,
// called per block at src/compression/compress.cu:259) is computed tile-
// parallel: every 4096-byte tile yields the CRC register of its bytes
// started from zero, and tiles are folded with GF(2) shifts
// (reg(A||B) = reg(A) * x^(8|B|) + reg0(B) mod P).
//
// RLE1 (reference: src/compression/rle1.cu:13-113) is restated per byte
// position: with the run [a, e) that contains position i and j = i - a,
// position i emits 0, 1 or 2 bytes (see rle1_emit), so the output is a
// stream compaction in input order: count per tile -> per-block scan ->
// write.  Run ends that cross tiles are found from per-tile boundary
// summaries instead of scanning bytes.
#include "pipeline.h"

namespace bzg {

struct Pow8Table {
  uint32_t p[40];
  constexpr Pow8Table() : p() {
    uint32_t b = 0x100u;  // x^8
    for (int k = 0; k < 40; k++) {
      p[k] = b;
      b = gf2_mulmod(b, b);
    }
  }
};

__constant__ CrcTable c_crc = CrcTable();
__constant__ Pow8Table c_pow8 = Pow8Table();

// x^(8*len) mod P from the constant table
__device__ inline uint32_t xpow8(uint32_t len) {
  uint32_t r = 1u;
  int k = 0;
  while (len) {
    if (len & 1u) r = gf2_mulmod(r, c_pow8.p[k]);
    len >>= 1;
    k++;
  }
  return r;
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
  for (int i = threadIdx.x; i < 6 * 260; i += blockDim.x) d.rfreq[(int64_t)b * 6 * 260 + i] = 0;
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

// Tile pass 1: CRC register of the tile (from zero) and the first/last run
// boundary (block-relative) of the tile.
__global__ void __launch_bounds__(256) k_front1(Dev d) {
  __shared__ uint32_t tab[256];
  __shared__ uint32_t regs[256];
  __shared__ uint32_t lens[256];
  __shared__ uint8_t lastb[256];
  __shared__ int32_t red[8];
  const int tid = threadIdx.x;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tn_blk[t];
  const BlockDesc D = d.desc[b];
  const int k = t - D.tile_n0;
  const int start = k * kTile;
  const int len = min(kTile, D.n - start);
  const uint8_t* src = d.in + D.in_off + start;
  tab[tid] = c_crc.t[tid];
  uint8_t x[16];
  int c = load16(src, len, tid, x);
  lastb[tid] = c > 0 ? x[c - 1] : 0;
  __syncthreads();
  uint32_t reg = 0;
  for (int q = 0; q < c; q++) reg = (reg << 8) ^ tab[(reg >> 24) ^ x[q]];
  regs[tid] = reg;
  lens[tid] = (uint32_t)c;
  // run boundaries
  int fb = 0x7fffffff, lb = -1;
  if (c > 0) {
    int p0 = start + tid * 16;
    uint8_t prev = 0;
    bool has_prev = false;
    if (tid > 0) { prev = lastb[tid - 1]; has_prev = true; }
    else if (start > 0) { prev = src[-1]; has_prev = true; }
    for (int q = 0; q < c; q++) {
      bool bnd = (q == 0) ? (!has_prev || x[0] != prev) : (x[q] != x[q - 1]);
      if (bnd) { int p = p0 + q; fb = min(fb, p); lb = max(lb, p); }
    }
  }
  __syncthreads();
  // tree-combine CRC registers: left * x^(8*len(right)) ^ right
  for (int s = 1; s < 256; s <<= 1) {
    if ((tid & (2 * s - 1)) == 0) {
      uint32_t rl = lens[tid + s];
      if (rl) regs[tid] = gf2_mulmod(regs[tid], xpow8(rl)) ^ regs[tid + s];
      lens[tid] += rl;
    }
    __syncthreads();
  }
  fb = wave_min(fb);
  lb = wave_max(lb);
  if (lane_id() == 0) { red[tid >> 6] = fb; red[4 + (tid >> 6)] = lb; }
  __syncthreads();
  if (tid == 0) {
    int f = red[0], l = red[4];
    for (int w = 1; w < 4; w++) { f = min(f, red[w]); l = max(l, red[4 + w]); }
    d.tn_crc[t] = regs[0];
    d.tn_fb[t] = f;
    d.tn_lb[t] = l;
  }
}

// Fold tile CRC registers into the block CRC (one wave per block).
__global__ void __launch_bounds__(64) k_crc_final(Dev d) {
  const int b = blockIdx.x;
  const BlockDesc D = d.desc[b];
  const int nt = tiles_for(D.n);
  const int l = lane_id();
  const int per = (nt + 63) / 64;
  int t0 = l * per, t1 = min(nt, t0 + per);
  uint32_t reg = 0, len = 0;
  for (int t = t0; t < t1; t++) {
    uint32_t tl = (uint32_t)min(kTile, D.n - t * kTile);
    reg = gf2_mulmod(reg, xpow8(tl)) ^ d.tn_crc[D.tile_n0 + t];
    len += tl;
  }
  for (int s = 1; s < 64; s <<= 1) {
    uint32_t r2 = __shfl_down(reg, s, 64);
    uint32_t l2 = __shfl_down(len, s, 64);
    if ((l & (2 * s - 1)) == 0 && l + s < 64) {
      if (l2) reg = gf2_mulmod(reg, xpow8(l2)) ^ r2;
      len += l2;
    }
  }
  if (l == 0) {
    uint32_t full = gf2_mulmod(0xFFFFFFFFu, xpow8((uint32_t)D.n)) ^ reg;
    d.st[b].crc = ~full;
  }
}

// Bytes emitted by raw position i of a run [a, e) of byte x
// (reference semantics: src/compression/rle1.cu:13-58).
__device__ inline int rle1_emit(int i, int a, int e, uint8_t x, uint8_t& v0, uint8_t& v1) {
  int c = e - a, j = i - a;
  int q = j / 259, r = j - 259 * q;
  int full = c / 259, rem = c - 259 * full;
  v0 = x;
  v1 = 0;
  if (q < full) {
    if (r < 4) return 1;
    if (r == 4) { v0 = 255; return 1; }
    return 0;
  }
  if (rem < 4) return r < rem ? 1 : 0;
  if (r < 3) return 1;
  if (r == 3) return rem == 4 ? 2 : 1;
  if (r == 4) { v0 = (uint8_t)(rem - 4); return 1; }
  return 0;
}

// Shared body of the RLE1 count and write passes.
template <bool WRITE>
__device__ inline void rle1_tile(const Dev& d) {
  __shared__ uint8_t lastb[256];
  __shared__ int32_t tmp[8];
  __shared__ int32_t carry[2];
  __shared__ uint32_t pres[8];
  const int tid = threadIdx.x;
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tn_blk[t];
  const BlockDesc D = d.desc[b];
  const int k = t - D.tile_n0;
  const int nt = tiles_for(D.n);
  const int start = k * kTile;
  const int len = min(kTile, D.n - start);
  const uint8_t* src = d.in + D.in_off + start;
  if (WRITE && tid < 8) pres[tid] = 0;
  // carries: last boundary before the tile (prefix max of tile lb),
  // first boundary at/after the tile end (suffix min of tile fb)
  {
    int lbv = -1;
    for (int j = k - 1 - tid; j >= 0; j -= 256) {
      int v = d.tn_lb[D.tile_n0 + j];
      if (v >= 0) { lbv = v; break; }
    }
    int fbv = 0x7fffffff;
    for (int j = k + 1 + tid; j < nt; j += 256) {
      int v = d.tn_fb[D.tile_n0 + j];
      if (v != 0x7fffffff) { fbv = v; break; }
    }
    lbv = wave_max(lbv);
    fbv = wave_min(fbv);
    if (lane_id() == 0) { tmp[tid >> 6] = lbv; tmp[4 + (tid >> 6)] = fbv; }
    __syncthreads();
    if (tid == 0) {
      int a = tmp[0], f = tmp[4];
      for (int w = 1; w < 4; w++) { a = max(a, tmp[w]); f = min(f, tmp[4 + w]); }
      carry[0] = a;
      carry[1] = f == 0x7fffffff ? D.n : f;
    }
    __syncthreads();
  }
  uint8_t x[16];
  int c = load16(src, len, tid, x);
  lastb[tid] = c > 0 ? x[c - 1] : 0;
  __syncthreads();
  const int p0 = start + tid * 16;
  uint8_t prev = 0;
  bool has_prev = false;
  if (tid > 0) { prev = lastb[tid - 1]; has_prev = true; }
  else if (start > 0) { prev = src[-1]; has_prev = true; }
  uint32_t bmask = 0;
  int lastB = -1, firstB = 0x7fffffff;
  for (int q = 0; q < c; q++) {
    bool bnd = (q == 0) ? (!has_prev || x[0] != prev) : (x[q] != x[q - 1]);
    if (bnd) {
      bmask |= 1u << q;
      lastB = p0 + q;
      if (firstB == 0x7fffffff) firstB = p0 + q;
    }
  }
  // exclusive prefix max of lastB over threads, exclusive suffix min of firstB
  int incl_a = wg_incl_max<256, int>(lastB, tmp, carry[0]);
  int incl_e = wg_rev_incl_min<256, int>(firstB, tmp, carry[1]);
  __shared__ int32_t xa[256], xe[256];
  xa[tid] = incl_a;
  xe[tid] = incl_e;
  __syncthreads();
  int ap = tid > 0 ? xa[tid - 1] : carry[0];
  int es = tid < 255 ? xe[tid + 1] : carry[1];
  // per-position runs
  int a_of[16], e_of[16];
  {
    int run = ap;
    for (int q = 0; q < 16; q++) {
      if (bmask & (1u << q)) run = p0 + q;
      a_of[q] = run;
    }
    int nb = es;
    for (int q = 15; q >= 0; q--) {
      e_of[q] = nb;
      if (bmask & (1u << q)) nb = p0 + q;
    }
  }
  int cnt = 0;
  uint8_t v0s[16], v1s[16], cs[16];
  for (int q = 0; q < 16; q++) {
    uint8_t v0 = 0, v1 = 0;
    int cc = q < c ? rle1_emit(p0 + q, a_of[q], e_of[q], x[q], v0, v1) : 0;
    cs[q] = (uint8_t)cc;
    v0s[q] = v0;
    v1s[q] = v1;
    cnt += cc;
  }
  int total;
  int excl = wg_excl_sum<256, int>(cnt, tmp, &total);
  if (!WRITE) {
    if (tid == 0) d.tn_cnt[t] = total;
    return;
  }
  uint8_t* dst = d.T + D.off + d.tn_off[t] + excl;
  int o = 0;
  for (int q = 0; q < 16; q++) {
    if (cs[q] >= 1) {
      dst[o++] = v0s[q];
      atomicOr(&pres[v0s[q] >> 5], 1u << (v0s[q] & 31));
    }
    if (cs[q] == 2) {
      dst[o++] = v1s[q];
      atomicOr(&pres[v1s[q] >> 5], 1u << (v1s[q] & 31));
    }
  }
  __syncthreads();
  if (tid < 8 && pres[tid]) atomicOr(&d.st[b].present[tid], pres[tid]);
}

__global__ void __launch_bounds__(256) k_rle1_count(Dev d) { rle1_tile<false>(d); }
__global__ void __launch_bounds__(256) k_rle1_write(Dev d) { rle1_tile<true>(d); }

// Per-block exclusive scan of tile counts.  which = 0: raw tiles -> st.m;
// which = 1: capacity tiles -> st.s (adds the end-of-block symbol).
__global__ void __launch_bounds__(256) k_tile_scan(Dev d, int which) {
  __shared__ int32_t tmp[8];
  const int b = blockIdx.x;
  const BlockDesc D = d.desc[b];
  const int t0 = which == 0 ? D.tile_n0 : D.tile_c0;
  const int nt = which == 0 ? tiles_for(D.n) : tiles_for(D.cap);
  const int32_t* cnt = which == 0 ? d.tn_cnt : d.tc_cnt;
  int32_t* off = which == 0 ? d.tn_off : d.tc_off;
  int run = 0;
  for (int base = 0; base < nt; base += 256) {
    int j = base + threadIdx.x;
    int v = j < nt ? cnt[t0 + j] : 0;
    int tot;
    int ex = wg_excl_sum<256, int>(v, tmp, &tot);
    if (j < nt) off[t0 + j] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) {
    if (which == 0) {
      d.st[b].m = run;
    } else {
      // end-of-block symbol = n_in_use + 1 (src/compression/rle2.cu:71-74)
      int eob = d.st[b].n_in_use + 1;
      d.SYM[D.off + run] = (uint16_t)eob;
      atomicAdd(&d.freq[(int64_t)b * 260 + eob], 1);
      d.st[b].s = run + 1;
    }
  }
}

void launch_block_init(const Dev& d, hipStream_t s) {
  k_block_init<<<d.nblk, 256, 0, s>>>(d);
}

void launch_crc_rle1(const Dev& d, hipStream_t s) {
  k_front1<<<d.ntile_n, 256, 0, s>>>(d);
  k_crc_final<<<d.nblk, 64, 0, s>>>(d);
  k_rle1_count<<<d.ntile_n, 256, 0, s>>>(d);
  k_tile_scan<<<d.nblk, 256, 0, s>>>(d, 0);
  k_rle1_write<<<d.ntile_n, 256, 0, s>>>(d);
}

void launch_tile_scan_c(const Dev& d, hipStream_t s) {
  k_tile_scan<<<d.nblk, 256, 0, s>>>(d, 1);
}

}  // namespace bzg
