// stage_aux.hip - small kernels that let a single stage run on its own
// (the per-stage entry points start from a stage's input, not from raw
// bytes, so the summaries an earlier stage would have produced are rebuilt).
#include "pipeline.h"

namespace bzg {

__global__ void __launch_bounds__(256) k_presence(Dev d) {
  __shared__ uint32_t pres[8];
  const int t = blockIdx.x;
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int m = d.st[b].m;
  const int start = (t - D.tile_c0) * kTile;
  if (threadIdx.x < 8) pres[threadIdx.x] = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < kTile && start + e < m; e += 256) {
    uint8_t c = d.T[D.off + start + e];
    atomicOr(&pres[c >> 5], 1u << (c & 31));
  }
  __syncthreads();
  if (threadIdx.x < 8 && pres[threadIdx.x]) atomicOr(&d.st[b].present[threadIdx.x], pres[threadIdx.x]);
}

__global__ void __launch_bounds__(256) k_fnz(Dev d) {
  __shared__ int32_t tmp[4];
  const int t = blockIdx.x;
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int m = d.st[b].m;
  const int start = (t - D.tile_c0) * kTile;
  int f = 0x7fffffff;
  for (int e = threadIdx.x; e < kTile && start + e < m; e += 256)
    if (d.MTF[D.off + start + e] != 0) f = min(f, start + e);
  f = wave_min(f);
  if (lane_id() == 0) tmp[threadIdx.x >> 6] = f;
  __syncthreads();
  if (threadIdx.x == 0) d.tc_fnz[t] = min(min(tmp[0], tmp[1]), min(tmp[2], tmp[3]));
}

__global__ void __launch_bounds__(256) k_sym_hist(Dev d) {
  __shared__ int32_t h[260];
  const int t = blockIdx.x;
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int s = d.st[b].s;
  const int start = (t - D.tile_c0) * kTile;
  for (int q = threadIdx.x; q < 260; q += 256) h[q] = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < kTile && start + e < s; e += 256) {
    uint16_t v = d.SYM[D.off + start + e];
    if (v < 260) atomicAdd(&h[v], 1);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 260; q += 256)
    if (h[q]) atomicAdd(&d.freq[(int64_t)b * 260 + q], h[q]);
}

// Range check of a caller's packing input (per-stage encode entry point):
// every symbol must index a row of kMaxAlpha code entries and every selector
// one of the n_groups tables; a violation sets the block's err flag.  Tile t
// checks symbols and selectors [start, start + kTile) of its block.
__global__ void __launch_bounds__(256) k_check_codes(Dev d, int n_groups) {
  const int t = blockIdx.x;
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int s = d.st[b].s, nsel = d.st[b].n_sel;
  const int start = (t - D.tile_c0) * kTile;
  int bad = 0;
  for (int e = threadIdx.x; e < kTile && start + e < s; e += 256)
    bad |= d.SYM[D.off + start + e] >= kMaxAlpha ? 1 : 0;
  for (int e = threadIdx.x; e < kTile && start + e < nsel; e += 256)
    bad |= d.sel[D.sel_off + start + e] >= n_groups ? 1 : 0;
  if (__syncthreads_or(bad) && threadIdx.x == 0) d.st[b].err = 1;
}

// Totals of a batch after its pipeline run (for the algorithmic-byte count
// of a benchmark): one workgroup sums the blocks' sizes.
__global__ void __launch_bounds__(256) k_block_totals(Dev d) {
  __shared__ int64_t part[4][kTotals];
  int64_t v[kTotals] = {0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < d.nblk; b += 256) {
    const BlockState& st = d.st[b];
    v[0] += 1;
    v[1] += d.desc[b].n;
    v[2] += st.m;
    v[3] += st.s;
    v[4] += st.bits;
    v[5] += (int64_t)st.m * st.rounds;
  }
#pragma unroll
  for (int k = 0; k < kTotals; k++) v[k] = wave_sum(v[k]);
  if (lane_id() == 0)
#pragma unroll
    for (int k = 0; k < kTotals; k++) part[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x < kTotals)
    d.totals[threadIdx.x] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

// A caller's suffix array -> internal entries (start | preceding byte << 24).
__global__ void __launch_bounds__(256) k_sa_attach(Dev d, const int32_t* sa) {
  const int t = blockIdx.x;
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int m = d.st[b].m;
  const int start = (t - D.tile_c0) * kTile;
  for (int e = threadIdx.x; e < kTile && start + e < m; e += 256) {
    const uint32_t v = (uint32_t)sa[start + e];
    const uint32_t c = v < (uint32_t)m ? d.T[D.off + (v == 0 ? (uint32_t)(m - 1) : v - 1)] : 0u;
    d.SA[D.off + start + e] = (v & kSAMask) | (c << 24);
  }
}

// Internal entries -> plain start indices.
__global__ void __launch_bounds__(256) k_sa_strip(Dev d, int32_t* sa) {
  const int t = blockIdx.x;
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int m = d.st[b].m;
  const int start = (t - D.tile_c0) * kTile;
  for (int e = threadIdx.x; e < kTile && start + e < m; e += 256)
    sa[start + e] = (int32_t)(d.SA[D.off + start + e] & kSAMask);
}

void launch_sa_attach(const Dev& d, const int32_t* sa, hipStream_t s) { k_sa_attach<<<d.ntile_c, 256, 0, s>>>(d, sa); }
void launch_sa_strip(const Dev& d, int32_t* sa, hipStream_t s) { k_sa_strip<<<d.ntile_c, 256, 0, s>>>(d, sa); }
void launch_presence(const Dev& d, hipStream_t s) { k_presence<<<d.ntile_c, 256, 0, s>>>(d); }
void launch_fnz(const Dev& d, hipStream_t s) { k_fnz<<<d.ntile_c, 256, 0, s>>>(d); }
void launch_sym_hist(const Dev& d, hipStream_t s) { k_sym_hist<<<d.ntile_c, 256, 0, s>>>(d); }
void launch_block_totals(const Dev& d, hipStream_t s) { k_block_totals<<<1, 256, 0, s>>>(d); }
void launch_check_codes(const Dev& d, int n_groups, hipStream_t s) { k_check_codes<<<d.ntile_c, 256, 0, s>>>(d, n_groups); }

}  // namespace bzg
