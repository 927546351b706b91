// bwt.hip - BWT suffix array of the cyclic rotations of every RLE1 block in
// a batch, by prefix doubling with early exit.
//
// Reference: src/compression/fbwt.cu:13-95 sorts ALL m (rank[SA[i]],
// rank[(SA[i]+k) mod m]) pairs with a 64-bit-key radix sort in each of
// ceil(log2 m) rounds.  Here, after an initial sort on the first 4 bytes,
// a round only touches rotations whose group (equal prefix so far) is not
// yet a singleton, and sorts each such group by one 21-bit key:
//     key(i) = RK[(SA[i] + P) mod m]
// where RK[p] is the position of the first row of p's group (the group
// "head") and P is the prefix length the groups currently represent.  A
// group whose block already has P >= m holds identical rotations; it is
// sorted by start index instead (the reference's stable sort from the
// identity gives the same order, tests/compression/fbwt.cu:22).
//
// Work per round (every kernel has a fixed, persistent grid and takes its
// counts from device memory, so rounds are enqueued, or replayed from a
// captured graph, without any host readback; a round after the last one
// finds an empty group list and does nothing):
//   plan      fixed-grid scans over the group list: prefix sums of small
//             group lengths (packed 2048-element LDS sorts) and of large
//             group tile counts; the round counter advances.
//   keys      gather key(i) for every row of every listed group.
//   small     one workgroup per pack: LDS radix sort of (group, key)
//             composite keys, new heads by a workgroup max-scan, RK/SA
//             written, subgroups of size > 1 appended to the next list.
//   large     segmented LSD radix sort over 4096-element tiles of the large
//             groups (histogram -> per-group scan -> stable scatter), then
//             boundary summaries and heads, as above.
#include "pipeline.h"

#include <algorithm>

namespace bzg {

// -------------------------------------------------------------- helpers
// Stable rank of this lane's digit among the items this wave has ranked so
// far; cnt = this wave's 256 counters in LDS.
__device__ inline uint32_t wave_rank_digit(uint32_t digit, bool valid, uint32_t* cnt) {
  const uint64_t peers = wave_match_bits<8>(digit, ballot(valid));
  uint32_t r = (uint32_t)mbcnt(peers);
  uint32_t base = 0;
  if (valid) {
    base = cnt[digit];
    int leader = __builtin_ctzll(peers);
    if (lane_id() == leader) cnt[digit] = base + (uint32_t)__popcll(peers);
  }
  return base + r;
}

__device__ inline int bits_for(uint32_t v) {  // bits to hold 0..v
  return v == 0 ? 0 : 32 - __clz(v);
}

// Wave-uniform reads.  Inside a persistent kernel's item loop the compiler
// cannot prove that the item tables (group list, tile maps, block
// descriptors) are unchanged by the earlier items' stores, so it loads them
// with vector loads and treats the values as per-lane: every address derived
// from them would become per-lane arithmetic.  readfirstlane puts the
// (uniform) values back in scalar registers.
__device__ inline int u32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t u64(uint64_t v) {
  return ((uint64_t)(uint32_t)u32((int)(v >> 32)) << 32) | (uint32_t)u32((int)(uint32_t)v);
}
__device__ inline int ctr_get(const Dev& d, int slot) { return u32(d.ctr[slot]); }
__device__ inline Seg seg_get(const Seg* p) {
  const Seg g = *p;
  return Seg{u32(g.blk), u32(g.start), u32(g.len), 0};
}

// The thread index as a value the compiler cannot hoist out of a persistent
// kernel's item loop: without it, every per-lane address offset of an item
// (sixteen per load sequence) is computed once before the loop and kept
// live across all items, which costs registers and occupancy.  The empty
// asm only marks the register as changed.
__device__ inline int loop_tid() {
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// ------------------------------------------------------------------ init
__global__ void __launch_bounds__(256) k_bwt_init(Dev d) {
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int m = d.st[b].m;
  const int k = t - D.tile_c0;
  // SA entries carry the row's BWT byte T[(i-1) mod m] in their top 8 bits;
  // KEY gets the round-0 key, the rotation's first 4 bytes.  A thread
  // handles four runs of 4 rows, run g at row k*kTile + 1024*g + 4*tid, so a
  // wave's 16-byte stores cover 1 KiB contiguously.
  const uint8_t* T = d.T + D.off;
  uint32_t* SA = d.SA + D.off;
  uint32_t* KEY = d.KEY + D.off;
  const int lane = lane_id();
  // a block of more than kHalf rows is one large group in round 0, whose
  // first radix pass makes these entries from the text itself
  const bool gen = m > kHalf;
#pragma unroll
  for (int g = 0; g < 4 && !gen; g++) {
    const int r0 = k * kTile + 1024 * g + 4 * (int)threadIdx.x;
    // rows r0..r0+3 need bytes r0-1 .. r0+6; the whole wave takes the fast
    // path when its last lane's 8 bytes lie inside the block
    const bool fast = k * kTile + 1024 * g + 4 * (int)(threadIdx.x | 63) + 8 <= m;
    if (fast) {
      const uint32_t w0 = *reinterpret_cast<const uint32_t*>(T + r0);
      const uint32_t w1 = *reinterpret_cast<const uint32_t*>(T + r0 + 4);
      // byte r0-1: the lane below's top byte; lane 0 reads it (row 0 wraps)
      uint32_t below = (uint32_t)__shfl_up((int)w0, 1, 64);
      uint32_t prev = below >> 24;
      if (lane == 0) prev = T[r0 == 0 ? m - 1 : r0 - 1];
      const uint64_t x = ((uint64_t)w1 << 32) | w0;  // bytes r0..r0+7, little-endian
      uint32_t o[4], kk[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        o[r] = (uint32_t)(r0 + r) | ((r == 0 ? prev : (uint32_t)(x >> (8 * r - 8)) & 255u) << 24);
        kk[r] = __builtin_bswap32((uint32_t)(x >> (8 * r)));
      }
      *reinterpret_cast<uint4*>(SA + r0) = make_uint4(o[0], o[1], o[2], o[3]);
      *reinterpret_cast<uint4*>(KEY + r0) = make_uint4(kk[0], kk[1], kk[2], kk[3]);
    } else {
      for (int i = r0; i < min(r0 + 4, m); i++) {
        SA[i] = (uint32_t)i | ((uint32_t)T[i == 0 ? m - 1 : i - 1] << 24);
        uint32_t kk = 0;
        int p = i;
        for (int q = 0; q < 4; q++) {
          kk = (kk << 8) | T[p];
          p = p + 1 == m ? 0 : p + 1;
        }
        KEY[i] = kk;
      }
    }
  }
  if (t == 0 && threadIdx.x == 0) {
    d.ctr[CTR_ROUND] = -1;
    d.ctr[CTR_KEPT] = 0;
    d.ctr[CTR_RAN] = 0;
  }
  if (k == 0 && threadIdx.x == 0) {
    d.st[b].rk = 0;
    d.st[b].rk_rows = 0;
    if (m > 1) {
      int idx = atomicAdd(&d.ctr[CTR_LIST0], 1);
      d.seg[0][idx] = Seg{b, 0, m, 0};
      d.st[b].rounds = 1;  // round 0 touches every block with m > 1
    } else {
      d.RK[D.off] = 0;
    }
  }
}

// ------------------------------------------------------------------ plan
struct PlanVals { int a, b, c, e; };
__device__ inline PlanVals plan_vals(const Seg& g) {
  bool small = g.len <= kHalf;
  return PlanVals{small ? g.len : 0, small ? 0 : tiles_for(g.len), small ? 0 : 1, small ? 0 : g.len};
}

// the group list a round reads: round r sorts list r & 1 and fills the other
__global__ void __launch_bounds__(256) k_plan_scan1(Dev d) {
  __shared__ int32_t tmp[4][8];
  const int cur = (ctr_get(d, CTR_ROUND) + 1) & 1;  // the round being planned
  const int n = ctr_get(d, cur);
  const int per = (n + kScanWG - 1) / kScanWG;
  const int i0 = blockIdx.x * per, i1 = min(n, i0 + per);
  int sa = 0, sb = 0, sc = 0, se = 0;
#pragma unroll 8
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    PlanVals v = plan_vals(d.seg[cur][i]);
    sa += v.a; sb += v.b; sc += v.c; se += v.e;
  }
  sa = wave_sum(sa); sb = wave_sum(sb); sc = wave_sum(sc); se = wave_sum(se);
  if (lane_id() == 0) {
    tmp[0][threadIdx.x >> 6] = sa; tmp[1][threadIdx.x >> 6] = sb;
    tmp[2][threadIdx.x >> 6] = sc; tmp[3][threadIdx.x >> 6] = se;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    int s = 0;
    for (int w = 0; w < 4; w++) s += tmp[threadIdx.x][w];
    d.scan_part[threadIdx.x * kScanWG + blockIdx.x] = s;
  }
}

// Totals of the round's plan; advances the round counter and records the
// rows the round sorts.  A list still holding groups after the last round
// the sort allows is an error (the round is then emptied).
__global__ void __launch_bounds__(kScanWG) k_plan_scan2(Dev d) {
  __shared__ int32_t tmp[kScanWG / 64 + 1];
  const int r = d.ctr[CTR_ROUND] + 1;
  const int cur = r & 1;
  int tots[4];
  for (int q = 0; q < 4; q++) {
    int v = d.scan_part[q * kScanWG + threadIdx.x];
    int tot;
    int ex = wg_excl_sum<kScanWG, int>(v, tmp, &tot);
    d.scan_part[q * kScanWG + threadIdx.x] = ex;
    tots[q] = tot;
  }
  if (threadIdx.x == 0) {
    int nseg = d.ctr[cur];
    if (nseg > 0 && r >= kMaxSortRounds - 1) {
      d.ctr[CTR_ERR] = 1;
      nseg = 0;
      tots[0] = tots[1] = tots[2] = tots[3] = 0;
    }
    d.ctr[CTR_SMALL] = tots[0];
    d.ctr[CTR_LTILES] = tots[1];
    d.ctr[CTR_NLARGE] = tots[2];
    d.ctr[CTR_LELEMS] = tots[3];
    d.ctr[CTR_NSEG] = nseg;
    d.ctr[CTR_ROUND] = r;
    d.ctr[CTR_RKN] = 0;  // blocks flagged for rank building count again
    d.ctr[CTR_RPK] = 0;
    if (nseg > 0) d.ctr[CTR_RAN] = r + 1;
    if (r < kMaxSortRounds) {  // rows in small groups, then in large groups
      d.round_rows[r] = nseg > 0 ? tots[0] : 0;
      d.round_rows[kMaxSortRounds + r] = nseg > 0 ? tots[3] : 0;
    }
  }
}

// Per group: exclusive prefixes of the small-group rows, large-group tiles
// and large-group count.  Each step takes 4096 groups, group base + 256q +
// tid in thread tid's slot q (coalesced, all loads in flight); the 64 (slot,
// wave) runs are in group order, so wave scans, a scan over the runs' totals
// by one wave, and the carried totals give every prefix.
__global__ void __launch_bounds__(256) k_plan_scan3(Dev d) {
  constexpr int kPer = 16;
  __shared__ int32_t tw[3][64];
  __shared__ int32_t tt[3];
  const int cur = ctr_get(d, CTR_ROUND) & 1;
  const int n = ctr_get(d, CTR_NSEG);
  const int per = (n + kScanWG - 1) / kScanWG;
  const int i0 = blockIdx.x * per, i1 = min(n, i0 + per);
  const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
  int ra = d.scan_part[0 * kScanWG + blockIdx.x];
  int rb = d.scan_part[1 * kScanWG + blockIdx.x];
  int rc = d.scan_part[2 * kScanWG + blockIdx.x];
  for (int base = i0; base < i1; base += 256 * kPer) {
    PlanVals v[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int i = base + 256 * q + tid;
      v[q] = i < i1 ? plan_vals(d.seg[cur][i]) : PlanVals{0, 0, 0, 0};
    }
    int xa[kPer], xb[kPer], xc[kPer];  // exclusive within the run
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int ia = wave_incl_sum(v[q].a), ib = wave_incl_sum(v[q].b), ic = wave_incl_sum(v[q].c);
      xa[q] = ia - v[q].a; xb[q] = ib - v[q].b; xc[q] = ic - v[q].c;
      if (lane == 63) { tw[0][4 * q + wv] = ia; tw[1][4 * q + wv] = ib; tw[2][4 * q + wv] = ic; }
    }
    __syncthreads();
    if (wv == 0) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int x = tw[k][lane];
        const int inc = wave_incl_sum(x);
        tw[k][lane] = inc - x;
        if (lane == 63) tt[k] = inc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int i = base + 256 * q + tid;
      if (i < i1) {
        const int j = 4 * q + wv;
        d.pa[i] = ra + tw[0][j] + xa[q];
        d.pb[i] = rb + tw[1][j] + xb[q];
        d.pc[i] = rc + tw[2][j] + xc[q];
      }
    }
    ra += tt[0]; rb += tt[1]; rc += tt[2];
    __syncthreads();
  }
}

void launch_bwt_plan(const Dev& d, hipStream_t s) {
  k_plan_scan1<<<kScanWG, 256, 0, s>>>(d);
  k_plan_scan2<<<1, kScanWG, 0, s>>>(d);
  k_plan_scan3<<<kScanWG, 256, 0, s>>>(d);
}

__device__ inline int round_packs(const Dev& d) { return (ctr_get(d, CTR_SMALL) + kHalf - 1) / kHalf; }

// packs whose groups are all at most this long are sorted by counting
constexpr int kCountMax = 64;
// round 1 builds the ranks of a block with more than 1/kRkShare of its rows
// in packs sorted by composite keys (k_pack_flags)
constexpr int kRkShare = 8;

// Ranks are kept only for blocks that need them.  Round 0 sorts by the text
// and rounds 1-2 sort the packs of small groups (at most kCountMax rows per
// group) by text keys too, so a block needs its ranks only when one of its
// groups takes the radix paths (more than kCountMax rows: a pack sorted by
// composite keys, or the large-group tiles), when it shares such a pack, or
// from round 3 on.  Such a block's ranks are then built from its suffix
// array and the round's group list (rank = the row's group start, or the row
// itself outside every group) and kept up to date by the later rounds.
__device__ inline void rk_need(const Dev& d, int b) {
  if (d.st[b].rk == 0 && atomicCAS(&d.st[b].rk, 0, 1) == 0) d.rk_list[atomicAdd(&d.ctr[CTR_RKN], 1)] = b;
}
// Work items of the blocks being built: (list entry, capacity tile of that
// block); items past a block's last tile return -1.
__device__ inline int rk_items(const Dev& d) { return ctr_get(d, CTR_RKN) * d.tiles_max; }
__device__ inline int rk_item_tile(const Dev& d, int item) {
  const int b = u32(d.rk_list[item / d.tiles_max]), k = item % d.tiles_max;
  const BlockDesc& D = d.desc[b];
  return k < tiles_for(u32(D.cap)) ? u32(D.tile_c0) + k : -1;
}
__device__ inline bool rk_kept(const Dev& d, int b) { return d.st[b].rk == 2; }

// pack k = small groups whose pa lies in [k*kHalf, (k+1)*kHalf)
__global__ void __launch_bounds__(256) k_mk_packs(Dev d) {
  const int npacks = round_packs(d), nseg = ctr_get(d, CTR_NSEG);
  for (int k = blockIdx.x * 256 + threadIdx.x; k <= npacks; k += gridDim.x * 256) {
    if (k == npacks) {
      d.packs[k] = nseg;
      continue;
    }
    int target = k * kHalf;
    int lo = 0, hi = nseg;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (d.pa[mid] < target) lo = mid + 1; else hi = mid;
    }
    d.packs[k] = lo;
  }
}

__global__ void __launch_bounds__(256) k_mk_large(Dev d) {
  const int round = ctr_get(d, CTR_ROUND), cur = round & 1, nseg = ctr_get(d, CTR_NSEG);
  if (blockIdx.x == 0 && threadIdx.x == 0) d.ctr[cur ^ 1] = 0;  // next list starts empty
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nseg; i += gridDim.x * 256) {
    Seg g = d.seg[cur][i];
    // (round 1 sorts every small group by text keys, round 2 the counting
    // ones, so ranks are needed for: round 1's large groups, round 2's
    // groups of more than kCountMax rows, every group from round 3 on)
    if (round >= 1 && (round >= 3 || g.len > (round == 1 ? kHalf : kCountMax))) rk_need(d, g.blk);
    if (g.len <= kHalf) continue;
    d.llist[d.pc[i]] = i;
    int nt = tiles_for(g.len), t0 = d.pb[i];
    for (int t = 0; t < nt; t++) {
      d.ltile_seg[t0 + t] = i;
      d.ltile_pos[t0 + t] = t * kTile;
    }
  }
}

// ------------------------------------------------------------ pack loading
struct PackLds {
  int32_t start[1024];
  int32_t blk[1024];
  int32_t lofs[1024];
  int32_t count;
  int32_t total;
  int32_t last_len;
  int32_t max_len;  // longest group
  int32_t no_rank;  // (with rank_state) a group's block keeps no ranks
  int32_t tmp[8];
};


// Loads the small groups of pack k into LDS (a pack holds <= 1023 groups:
// each has >= 2 rows and the pack <= kPack-1 rows).
__device__ inline void load_pack(const Dev& d, int cur, int k, PackLds& P, bool rank_state = false) {
  const int tid_ = loop_tid();
  const int i0 = u32(d.packs[k]), i1 = u32(d.packs[k + 1]);
  int count = 0;
  if (tid_ == 0) {
    P.max_len = 0;
    P.no_rank = 0;
  }
  for (int base = i0; base < i1; base += 256) {
    int i = base + tid_;
    Seg g{0, 0, 0, 0};
    bool small = false;
    if (i < i1) { g = d.seg[cur][i]; small = g.len <= kHalf; }
    int tot;
    int ex = wg_excl_sum<256, int>(small ? 1 : 0, P.tmp, &tot);
    if (small) {
      int idx = count + ex;
      P.start[idx] = g.start;
      P.blk[idx] = g.blk;
      P.lofs[idx] = d.pa[i];
      if (ex == tot - 1) P.last_len = g.len;
    }
    const int ml = wave_max(small ? g.len : 0);
    if ((tid_ & 63) == 0 && ml > 0) atomicMax(&P.max_len, ml);
    if (rank_state) {
      const bool nr = __ballot(small && d.st[g.blk].rk != 2) != 0;
      if ((tid_ & 63) == 0 && nr) atomicOr(&P.no_rank, 1);
    }
    count += tot;
    __syncthreads();
  }
  __syncthreads();
  if (count > 0) {
    const int base_pa = P.lofs[0];
    __syncthreads();
    for (int j = tid_; j < count; j += 256) P.lofs[j] -= base_pa;
    __syncthreads();
    if (tid_ == 0) P.total = P.lofs[count - 1] + P.last_len;
  }
  if (tid_ == 0) P.count = count;
  __syncthreads();
}

__device__ inline int find_seg(const PackLds& P, int e) {  // last j with lofs[j] <= e
  int lo = 0, hi = P.count - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (P.lofs[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------------- keys
// The rotation's 4 bytes at start q (cyclic, big-endian: the round-0 key of
// q), for a block of m > 4 bytes at T (4-byte aligned).
__device__ inline uint32_t text_key(const uint8_t* T, uint32_t q, uint32_t m) {
  const uint32_t w = q >> 2;
  uint32_t x;
  if (4 * w + 8 <= m) {  // both words inside the block
    const uint32_t* T32 = reinterpret_cast<const uint32_t*>(T);
    x = __builtin_amdgcn_alignbyte(T32[w + 1], T32[w], q & 3u);
  } else {
    x = 0;
    uint32_t p = q;
    for (int k = 0; k < 4; k++) {
      x |= (uint32_t)T[p] << (8 * k);
      p = p + 1 == m ? 0 : p + 1;
    }
  }
  return __builtin_bswap32(x);
}

// The rotation's 8 bytes at start q, big-endian, for a block of m > 8 bytes.
__device__ inline uint64_t text_key8(const uint8_t* T, uint32_t q, uint32_t m) {
  const uint32_t w = q >> 2;
  uint32_t lo, hi;  // bytes q..q+3 and q+4..q+7, little-endian
  if (4 * w + 12 <= m) {  // the three words inside the block
    const uint32_t* T32 = reinterpret_cast<const uint32_t*>(T);
    const uint32_t a = T32[w], b = T32[w + 1], c = T32[w + 2];
    lo = __builtin_amdgcn_alignbyte(b, a, q & 3u);
    hi = __builtin_amdgcn_alignbyte(c, b, q & 3u);
  } else {
    lo = hi = 0;
    uint32_t p = q;
    for (int k = 0; k < 8; k++) {
      const uint32_t v = (uint32_t)T[p] << (8 * (k & 3));
      if (k < 4) lo |= v; else hi |= v;
      p = p + 1 == m ? 0 : p + 1;
    }
  }
  return ((uint64_t)__builtin_bswap32(lo) << 32) | __builtin_bswap32(hi);
}

// Rounds r >= 1: key(i) = RK[(SA[i] + P) mod m] for every row of every
// listed group (P >= m: identical rotations, ordered by start index).
// Item w < npacks is a small-group pack, the others large-group tiles.
// In round 1 (P = 4) the ranks would be those of round 0, which order the
// rotations by their first 4 bytes, so the 4 text bytes at SA + 4 give the
// same order and the same ties; in round 2 (P = 8) the 8 text bytes at
// SA + 8 stand for the ranks of round 1 the same way.  Packs sorted by
// counting (groups of at most kCountMax rows, compared only within a group)
// take those, and since the text does not change during the round, the
// sort gathers them itself (`text_keys`): such packs have nothing to do
// here but mark their blocks' round count.  A block's text is a quarter of
// its rank array, so the gathers of its rows also share cache lines four
// times as often, and blocks whose groups all stay that small need no ranks
// at all.  In round 1 the packs sorted by composite keys take text keys
// too when a block of theirs keeps no ranks (32 key bits instead of the
// rank bits, in a 64-bit sort of their own).
__device__ inline bool text_keys(int P, const PackLds& L) {
  return L.max_len <= kCountMax ? (P == 4 || P == 8) : (P == 4 && L.no_rank);
}
// Items w < npacks are packs (the listed ones, rpk[w], with `listed`), the
// others large-group tiles.  (In rounds 1-2 the packs' blocks get their
// round count from the sort, which loads every pack.)
__device__ inline void keys_item(const Dev& d, int cur, int round, int P, int npacks, int w, PackLds& L,
                                 bool listed) {
  const int tid_ = loop_tid();
  if (w < npacks) {
    load_pack(d, cur, listed ? u32(d.rpk[w]) : w, L, P == 4);
    if (L.count == 0) return;
    const int total = L.total;
    if (!listed)
      for (int j = tid_; j < L.count; j += 256)
        if (j == 0 || L.blk[j] != L.blk[j - 1]) atomicMax(&d.st[L.blk[j]].rounds, round + 1);
    if (text_keys(P, L)) return;
    // rows' segment and block first (block data reloaded only when the block
    // changes), then all SA loads, then all rank loads, then the stores
    int jj[kPack / 256], mm[kPack / 256];
    uint64_t oo[kPack / 256];
    int cb = -1, cm = 0;
    uint64_t coff = 0;
#pragma unroll
    for (int q = 0; q < kPack / 256; q++) {
      const int e = tid_ + 256 * q;
      jj[q] = -1;
      mm[q] = 1;
      oo[q] = 0;
      if (e < total) {
        const int s = find_seg(L, e);
        const int b = L.blk[s];
        if (b != cb) {
          cb = b;
          coff = d.desc[b].off;
          cm = d.st[b].m;
        }
        jj[q] = L.start[s] + (e - L.lofs[s]);
        oo[q] = coff;
        mm[q] = cm;
      }
    }
    uint32_t v[kPack / 256];
#pragma unroll
    for (int q = 0; q < kPack / 256; q++) v[q] = jj[q] >= 0 ? d.SA[oo[q] + jj[q]] & kSAMask : 0u;
#pragma unroll
    for (int q = 0; q < kPack / 256; q++) {
      if (jj[q] >= 0 && P < mm[q]) {
        uint32_t p = v[q] + (uint32_t)P;
        if (p >= (uint32_t)mm[q]) p -= (uint32_t)mm[q];
        v[q] = d.RK[oo[q] + p];
      }
    }
#pragma unroll
    for (int q = 0; q < kPack / 256; q++)
      if (jj[q] >= 0) d.KEY[oo[q] + jj[q]] = v[q];
  } else {
    const int t = w - npacks;
    const int si = u32(d.ltile_seg[t]);
    const Seg g = seg_get(&d.seg[cur][si]);
    const int pos = u32(d.ltile_pos[t]);
    const uint64_t off = u64(d.desc[g.blk].off);
    const int m = u32(d.st[g.blk].m);
    if (pos == 0 && tid_ == 0) atomicMax(&d.st[g.blk].rounds, round + 1);
    const int n = min(kTile, g.len - pos);
    // all SA loads, then all rank loads, then the stores
    const int j0 = g.start + pos;
    uint32_t v[kTile / 256];
#pragma unroll
    for (int q = 0; q < kTile / 256; q++) {
      const int e = tid_ + 256 * q;
      v[q] = e < n ? d.SA[off + j0 + e] & kSAMask : 0u;
    }
    if (P < m) {
#pragma unroll
      for (int q = 0; q < kTile / 256; q++) {
        const int e = tid_ + 256 * q;
        if (e < n) {
          uint32_t p = v[q] + (uint32_t)P;
          if (p >= (uint32_t)m) p -= (uint32_t)m;
          v[q] = d.RK[off + p];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kTile / 256; q++) {
      const int e = tid_ + 256 * q;
      if (e < n) d.KEY[off + j0 + e] = v[q];
    }
  }
}

// (at most 64 VGPRs: eight workgroups per CU instead of six for the random
// rank gathers)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_keys(Dev d) {
  __shared__ PackLds L;
  const int round = ctr_get(d, CTR_ROUND), cur = round & 1, P = round_prefix(round);
  // rounds 1-2: only packs sorted by composite keys can need rank keys
  // (the others are sorted by counting, on text keys), and k_pack_flags
  // lists them
  const bool listed = round == 1 || round == 2;
  const int npacks = listed ? ctr_get(d, CTR_RPK) : round_packs(d);
  for (ItemLoop it(npacks + ctr_get(d, CTR_LTILES)); it.i < it.end; it.i += it.step) {
    __syncthreads();  // the previous item's pack is no longer read
    keys_item(d, cur, round, P, npacks, it.i, L, listed);
  }
}

// Appends groups to the next list: each thread holds `cnt` groups in g[].
template <int N>
__device__ inline void emit_groups_n(const Dev& d, int nxt, int cnt, const Seg* g, int32_t* tmp, int32_t* s_base) {
  int tot;
  int ex = wg_excl_sum<256, int>(cnt, tmp, &tot);
  if (threadIdx.x == 0 && tot > 0) *s_base = atomicAdd(&d.ctr[nxt], tot);
  __syncthreads();
  if (tot > 0) {
    const int base = *s_base + ex;
#pragma unroll
    for (int q = 0; q < N; q++) {
      if (q < cnt) {
        int idx = base + q;
        if (idx < d.seg_cap) d.seg[nxt][idx] = g[q];
        else d.ctr[CTR_ERR] = 1;
      }
    }
  }
  __syncthreads();
}

// Appends groups to the next list: each thread has 0/1 group to emit.
__device__ inline void emit_groups(const Dev& d, int nxt, bool has, Seg g, int32_t* tmp, int32_t* s_base) {
  int tot;
  int ex = wg_excl_sum<256, int>(has ? 1 : 0, tmp, &tot);
  if (threadIdx.x == 0 && tot > 0) *s_base = atomicAdd(&d.ctr[nxt], tot);
  __syncthreads();
  if (has) {
    int idx = *s_base + ex;
    if (idx < d.seg_cap) d.seg[nxt][idx] = g;
    else d.ctr[CTR_ERR] = 1;
  }
  __syncthreads();
}

// ----------------------------------------------------------- small groups
// One workgroup sorts a pack of <= kPack rows in LDS.  KT holds the
// composite key (group in pack << keybits | key): 32 bits whenever they fit.
// Every row of a group already has the group's first row as its rank (from
// the round that formed the group), and the rows of its first subgroup keep
// it: after round 0 (which ranks every row first) only the other rows'
// ranks are written (`keep_first`).
template <typename KT>
struct SortLds {
  PackLds L;
  KT K0[kPack], K1[kPack];
  uint32_t V0[kPack], V1[kPack];
  uint32_t cnt[4][256];
  int32_t xs[256];
  int32_t s_base;
  int32_t tmp[8];
};

// TEXT: the round may take keys from the text (round 1); the other rounds
// run a copy without that path (both copies share the LDS in S).
template <typename KT, bool TEXT>
// only: 0 every pack, 2 the packs sorted by composite text keys, 1 the
// others (round 1 splits them between two launches)
__device__ inline void sort_small_body(const Dev& d, int cur, bool keep_first, int P, int keybits, int k,
                                       SortLds<KT>& S, int only) {
  const int tid_ = loop_tid();
  PackLds& L = S.L;
  KT* K0 = S.K0;
  KT* K1 = S.K1;
  uint32_t* V0 = S.V0;
  uint32_t* V1 = S.V1;
  uint32_t (*cnt)[256] = S.cnt;
  int32_t* xs = S.xs;
  int32_t& s_base = S.s_base;
  int32_t* tmp = S.tmp;
  const int tid = tid_, wv = tid >> 6, lane = (tid_ & 63);
  const int nxt = cur ^ 1;
  load_pack(d, cur, k, L, TEXT && P == 4);
  if (L.count == 0) return;
  const int total = L.total;
  if (TEXT && only != 2)  // rounds 1-2 (P = 4, 8): the pack's blocks ran round P / 4
    for (int j = tid_; j < L.count; j += 256)
      if (j == 0 || L.blk[j] != L.blk[j - 1]) atomicMax(&d.st[L.blk[j]].rounds, P / 4 + 1);
  const bool counting = L.max_len <= kCountMax;
  // rounds 1-2: counting-sorted packs gather their keys from the text (in
  // the key fill below), 4 bytes at SA + 4 in round 1, 8 bytes at SA + 8 in
  // round 2 (the high half in V1, which the counting sort fills only after
  // its compares; K1 holds the rows' group indices)
  const bool tk = TEXT && text_keys(P, L);
  if (only != 0 && only != (tk && !counting ? 2 : 1)) return;
  const bool wide = tk && P == 8;
  // group of every row: mark group starts, then a running max over rows
  uint16_t* sid = reinterpret_cast<uint16_t*>(K1);  // free until the first pass
  for (int e = tid; e < kPack; e += 256) sid[e] = 0;
  __syncthreads();
  for (int j = tid; j < L.count; j += 256) sid[L.lofs[j]] = (uint16_t)j;
  __syncthreads();
  {
    const int e0 = tid * (kPack / 256);
    int mx = 0;
#pragma unroll
    for (int q = 0; q < kPack / 256; q++) mx = max(mx, (int)sid[e0 + q]);
    const int incl = wg_incl_max<256, int>(mx, tmp, 0);
    xs[tid] = incl;
    __syncthreads();
    int run = tid > 0 ? xs[tid - 1] : 0;
#pragma unroll
    for (int q = 0; q < kPack / 256; q++) {
      run = max(run, (int)sid[e0 + q]);
      sid[e0 + q] = (uint16_t)run;
    }
  }
  __syncthreads();
  // composite keys and SA values (a pack's groups come from few blocks: the
  // block offset is reloaded only when the block changes, and the rows'
  // loads are all in flight before the LDS writes)
  {
    int cb = -1, cm = 0;
    uint64_t coff = 0;
    uint32_t kv[kPack / 256], sv[kPack / 256];
    uint64_t oo[kPack / 256];
    int mm[kPack / 256];
#pragma unroll
    for (int q = 0; q < kPack / 256; q++) {
      const int e = tid + 256 * q;
      kv[q] = 0;
      sv[q] = 0;
      oo[q] = 0;
      mm[q] = 0;
      if (e < total) {
        const int sg = sid[e];
        const int b = L.blk[sg];
        if (b != cb) {
          cb = b;
          coff = d.desc[b].off;
          cm = tk ? d.st[b].m : 0;
        }
        const int j = L.start[sg] + (e - L.lofs[sg]);
        if (!tk) kv[q] = d.KEY[coff + j];
        sv[q] = d.SA[coff + j];
        oo[q] = coff;
        mm[q] = cm;
      }
    }
    uint32_t kh[kPack / 256];
#pragma unroll
    for (int q = 0; q < kPack / 256; q++) kh[q] = 0;
    if (tk) {
#pragma unroll
      for (int q = 0; q < kPack / 256; q++) {
        if (tid + 256 * q < total) {
          // key(i) = the P bytes at SA + P (the start itself when the
          // prefix covers the rotation: identical rotations by start)
          const uint32_t a = sv[q] & kSAMask, cmq = (uint32_t)mm[q];
          uint32_t p = a + (uint32_t)P;
          if (p >= cmq) p -= cmq;
          if (!wide) {
            kv[q] = 4 < mm[q] ? text_key(d.T + oo[q], p, cmq) : a;
          } else if (8 < mm[q]) {
            const uint64_t k8 = text_key8(d.T + oo[q], p, cmq);
            kv[q] = (uint32_t)k8;
            kh[q] = (uint32_t)(k8 >> 32);
          } else {
            kv[q] = a;
          }
        }
      }
    }
    // (the counting sort compares keys within a group only: the key alone,
    // which in rounds 1-2 may be 32 or 64 text bits)
#pragma unroll
    for (int q = 0; q < kPack / 256; q++) {
      const int e = tid + 256 * q;
      if (e < total) {
        if (counting && sizeof(KT) == 8) K0[e] = (KT)(((uint64_t)kh[q] << 32) | kv[q]);
        else K0[e] = counting ? (KT)kv[q] : (((KT)sid[e] << keybits) | (KT)kv[q]);
        if (wide && sizeof(KT) == 4) V1[e] = kh[q];
        V0[e] = sv[q];
      }
    }
  }
  // 64-bit keys of the counting sort: K0 alone, or V1 (high) and K0 (low)
  const bool split = wide && sizeof(KT) == 4;
  auto key_at = [&](int i) -> uint64_t {
    return split ? (((uint64_t)V1[i] << 32) | (uint32_t)K0[i]) : (uint64_t)K0[i];
  };
  __syncthreads();
  if (counting) {
    // Small groups only: every row's place in its group by counting (rows
    // with a smaller key, and equal keys before it); equal keys stay together
    // in input order and form the next round's subgroup.
    int cb = -1;
    bool ck = false;  // the row's block keeps ranks
    uint64_t coff = 0;
    uint32_t vres[kPack / 256];  // subgroup of each row, to V1 after the compares
#pragma unroll
    for (int r = 0; r < kPack / 256; r++) {
      const int e = tid + 256 * r;
      vres[r] = 0;
      if (e >= total) continue;
      const int sg = sid[e];
      const int gs = L.lofs[sg];
      const int ge = sg + 1 < L.count ? L.lofs[sg + 1] : total;
      int lt = 0, eq = 0, eqb = 0;
      if (split) {
        const uint64_t my = key_at(e);
        for (int j = gs; j < ge; j++) {
          const uint64_t kj = key_at(j);
          lt += kj < my ? 1 : 0;
          const int same = kj == my ? 1 : 0;
          eq += same;
          eqb += (j < e) ? same : 0;
        }
      } else {
        const KT my = K0[e];
        for (int j = gs; j < ge; j++) {
          const KT kj = K0[j];
          lt += kj < my ? 1 : 0;
          const int same = kj == my ? 1 : 0;
          eq += same;
          eqb += (j < e) ? same : 0;
        }
      }
      const int b = L.blk[sg];
      if (b != cb) {
        cb = b;
        coff = d.desc[b].off;
        ck = rk_kept(d, b);
      }
      const int st0 = L.start[sg];
      const uint32_t v = V0[e];
      d.SA[coff + st0 + lt + eqb] = v;
      if (ck && (lt != 0 || !keep_first)) d.RK[coff + (v & kSAMask)] = (uint32_t)(st0 + lt);
      vres[r] = (eq >= 2 && eqb == 0) ? (((uint32_t)lt << 16) | (uint32_t)eq) : 0u;
    }
    __syncthreads();  // every row's compares are done (V1 may hold key halves)
#pragma unroll
    for (int r = 0; r < kPack / 256; r++)
      if (tid + 256 * r < total) V1[tid + 256 * r] = vres[r];
    __syncthreads();
    // subgroups of >= 2 rows, emitted by their first row, in row order
    const int e0 = tid * (kPack / 256);
    int nout = 0;
#pragma unroll
    for (int q = 0; q < kPack / 256; q++)
      if (e0 + q < total && V1[e0 + q]) nout++;
    int tot;
    const int ex = wg_excl_sum<256, int>(nout, tmp, &tot);
    if (tid == 0 && tot > 0) s_base = atomicAdd(&d.ctr[nxt], tot);
    __syncthreads();
    if (nout) {
      int idx = s_base + ex;
#pragma unroll
      for (int q = 0; q < kPack / 256; q++) {
        const int e = e0 + q;
        const uint32_t w = e < total ? V1[e] : 0u;
        if (w) {
          const int sg = sid[e];
          const Seg ng{L.blk[sg], L.start[sg] + (int)(w >> 16), (int)(w & 0xFFFFu), 0};
          if (idx < d.seg_cap) d.seg[nxt][idx] = ng;
          else d.ctr[CTR_ERR] = 1;
          idx++;
        }
      }
    }
    return;
  }
  __syncthreads();
  const int nbits = keybits + bits_for((uint32_t)(L.count - 1));
  const int npass = (nbits + 7) / 8;
  KT* Ks = K0; KT* Kd = K1;
  uint32_t* Vs = V0; uint32_t* Vd = V1;
  for (int pass = 0; pass < npass; pass++) {
    const int shift = pass * 8;
    cnt[0][tid] = 0; cnt[1][tid] = 0; cnt[2][tid] = 0; cnt[3][tid] = 0;
    __syncthreads();
    uint32_t rk[8], dg[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int e = wv * 512 + q * 64 + lane;
      bool valid = e < total;
      uint32_t digit = valid ? (uint32_t)(Ks[e] >> shift) & 255u : 0u;
      dg[q] = digit;
      rk[q] = wave_rank_digit(digit, valid, cnt[wv]);
    }
    __syncthreads();
    {
      uint32_t c0 = cnt[0][tid], c1 = cnt[1][tid], c2 = cnt[2][tid], c3 = cnt[3][tid];
      int tot;
      int base = wg_excl_sum<256, int>((int)(c0 + c1 + c2 + c3), tmp, &tot);
      cnt[0][tid] = base;
      cnt[1][tid] = base + c0;
      cnt[2][tid] = base + c0 + c1;
      cnt[3][tid] = base + c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int e = wv * 512 + q * 64 + lane;
      if (e < total) {
        uint32_t dst = cnt[wv][dg[q]] + rk[q];
        Kd[dst] = Ks[e];
        Vd[dst] = Vs[e];
      }
    }
    __syncthreads();
    KT* tk = Ks; Ks = Kd; Kd = tk;
    uint32_t* tv = Vs; Vs = Vd; Vd = tv;
  }
  // heads: thread owns rows [8*tid, 8*tid+8)
  const int e0 = tid * 8;
  int lastB = -1;
  uint32_t bm = 0;
  KT kk[9];
#pragma unroll
  for (int q = 0; q < 9; q++) kk[q] = (e0 + q < total) ? Ks[e0 + q] : (KT)0;
  {
    KT prev = e0 > 0 ? Ks[e0 - 1] : (KT)0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int e = e0 + q;
      if (e < total && (e == 0 || kk[q] != prev)) { bm |= 1u << q; lastB = e; }
      prev = kk[q];
    }
  }
  int incl = wg_incl_max<256, int>(lastB, tmp, -1);
  xs[tid] = incl;
  __syncthreads();
  int run = tid > 0 ? xs[tid - 1] : -1;
  int head_of[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    if (bm & (1u << q)) run = e0 + q;
    head_of[q] = run;
  }
  uint32_t endm = 0;
  int nout = 0;
  int cb = -1;
  bool ck = false;  // the row's block keeps ranks
  uint64_t coff = 0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    int e = e0 + q;
    if (e < total) {
      int sg = (int)(kk[q] >> keybits);
      int b = L.blk[sg];
      if (b != cb) {
        cb = b;
        coff = d.desc[b].off;
        ck = rk_kept(d, b);
      }
      const uint64_t off = coff;
      int j = L.start[sg] + (e - L.lofs[sg]);
      int hj = L.start[sg] + (head_of[q] - L.lofs[sg]);
      uint32_t v = Vs[e];
      d.SA[off + j] = v;
      if (ck && (head_of[q] != L.lofs[sg] || !keep_first)) d.RK[off + (v & kSAMask)] = (uint32_t)hj;
      // a group of >= 2 rows ends here
      if (((e == total - 1) || (kk[q + 1] != kk[q])) && e > head_of[q]) {
        endm |= 1u << q;
        nout++;
      }
    }
  }
  int tot;
  const int ex = wg_excl_sum<256, int>(nout, tmp, &tot);
  if (tid == 0 && tot > 0) s_base = atomicAdd(&d.ctr[nxt], tot);
  __syncthreads();
  if (endm) {
    int idx = s_base + ex;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      if ((endm >> q) & 1u) {
        const int sg = (int)(kk[q] >> keybits);
        const Seg ng{L.blk[sg], L.start[sg] + (head_of[q] - L.lofs[sg]), e0 + q - head_of[q] + 1, 0};
        if (idx < d.seg_cap) d.seg[nxt][idx] = ng;
        else d.ctr[CTR_ERR] = 1;
        idx++;
      }
    }
  }
}

// Round 1 sorts the packs sorted by composite text keys (32 text bits under
// the group index) in a second, 64-bit launch (radix1), the others in the
// round's usual one, so that no block needs ranks in round 1 for its small
// groups.
template <typename KT>
__device__ inline void sort_small_loop(const Dev& d, int keybits, int radix1) {
  __shared__ SortLds<KT> S;
  const int round = ctr_get(d, CTR_ROUND), cur = round & 1;
  if (radix1 && round != 1) return;
  // (the second launch goes over the packs that k_pack_flags listed)
  for (ItemLoop it(radix1 ? ctr_get(d, CTR_RPK) : round_packs(d)); it.i < it.end; it.i += it.step) {
    __syncthreads();  // the previous pack's LDS is no longer read
    if (round == 1)
      sort_small_body<KT, true>(d, cur, true, 4, radix1 ? 32 : keybits, radix1 ? u32(d.rpk[it.i]) : it.i, S,
                                radix1 ? 2 : 1);
    else if (round == 2)
      sort_small_body<KT, true>(d, cur, true, 8, keybits, it.i, S, 0);
    else
      sort_small_body<KT, false>(d, cur, round > 0, 0, keybits, it.i, S, 0);
  }
}

__global__ void __launch_bounds__(256) k_sort_small(Dev d, int keybits) { sort_small_loop<uint32_t>(d, keybits, 0); }

__global__ void __launch_bounds__(256) k_sort_small64(Dev d, int keybits, int radix1) {
  sort_small_loop<uint64_t>(d, keybits, radix1);
}

// ----------------------------------------------------------- large groups
// Digit histogram of a large-group tile (per-wave counters, so that skewed
// digits collide only within a wave).
// With GEN (round 0's first pass, whose groups are whole blocks) the digit
// is read from the text: the key's low byte is T[(i+3) mod m].
template <bool GEN>
__device__ inline void lhist_item(const Dev& d, int cur, int shift, int src, int t) {
  const int tid_ = loop_tid();
  __shared__ uint32_t h[4][256];
  const int tid = tid_, wv = tid >> 6;
  const Seg g = seg_get(&d.seg[cur][u32(d.ltile_seg[t])]);
  const int pos = u32(d.ltile_pos[t]);
  const uint64_t off = u64(d.desc[g.blk].off);
  const uint32_t* K = (src ? d.KEY2 : d.KEY) + off + g.start + pos;
  h[0][tid] = 0; h[1][tid] = 0; h[2][tid] = 0; h[3][tid] = 0;
  __syncthreads();
  const int n = min(kTile, g.len - pos);
  uint32_t k[16];
  if (GEN) {
    // digits of rows pos .. pos+n-1 are the bytes at pos+3 .. pos+n+2: each
    // thread takes 16 aligned bytes at pos + 16*tid (the rows 3 before
    // them), thread 255 also the 3 bytes after the tile (mod m)
    const uint8_t* T = d.T + off;
    const int m = g.len;
    const int q0 = 16 * tid;
    uint32_t w4[4] = {0u, 0u, 0u, 0u};
    if (q0 + 16 <= n) {
      const uint4 v = *reinterpret_cast<const uint4*>(T + pos + q0);
      w4[0] = v.x; w4[1] = v.y; w4[2] = v.z; w4[3] = v.w;
    } else {
#pragma unroll
      for (int j = 0; j < 16; j++)
        if (q0 + j < n) w4[j >> 2] |= (uint32_t)T[pos + q0 + j] << (8 * (j & 3));
    }
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int e = q0 + j - 3;  // the row whose digit this byte is
      const uint32_t dg = (w4[j >> 2] >> (8 * (j & 3))) & 255u;
      k[j] = (q0 + j < n && e >= 0) ? dg : 0xFFFFFFFFu;
    }
    if (tid == 255) {
#pragma unroll
      for (int j = 0; j < 3; j++) {
        int i = pos + n + j;
        i = i >= m ? i - m : i;  // m > kHalf > 3
        if (n - 3 + j >= 0) atomicAdd(&h[wv][T[i]], 1u);  // row n-3+j
      }
    }
#pragma unroll
    for (int j = 0; j < 16; j++)
      if (k[j] != 0xFFFFFFFFu) atomicAdd(&h[wv][k[j]], 1u);
    __syncthreads();
    d.hist[(int64_t)t * 256 + tid] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
    return;
  } else {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = tid + 256 * q;
      k[q] = e < n ? K[e] : 0u;
    }
  }
#pragma unroll
  for (int q = 0; q < 16; q++)
    if (tid + 256 * q < n) atomicAdd(&h[wv][(k[q] >> shift) & 255u], 1u);
  __syncthreads();
  d.hist[(int64_t)t * 256 + tid] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
}

template <bool GEN>
__global__ void __launch_bounds__(256) k_lhist(Dev d, int shift, int src) {
  const int cur = ctr_get(d, CTR_ROUND) & 1;
  for (ItemLoop it(ctr_get(d, CTR_LTILES)); it.i < it.end; it.i += it.step) {
    __syncthreads();
    lhist_item<GEN>(d, cur, shift, src, it.i);
  }
}

// per large group: hist[t][digit] -> absolute destination base
// Per large group and digit: exclusive prefix of the digit's count over the
// group's tiles.  The first tile's prefix is always 0, so its entry holds
// the digit's global start instead (group start + rows of smaller digits);
// k_lscatter adds it back for the other tiles.
__device__ inline void lscan_item(const Dev& d, int cur, int li) {
  const int tid_ = loop_tid();
  __shared__ int32_t tmp[8];
  const int si = u32(d.llist[li]);
  const Seg g = seg_get(&d.seg[cur][si]);
  const int t0 = u32(d.pb[si]), nt = tiles_for(g.len);
  const int dg = tid_;
  uint32_t* hp = d.hist + (int64_t)t0 * 256 + dg;
  uint32_t run = 0;
  for (int t = 0; t < nt; t += 16) {
    uint32_t v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = t + q < nt ? hp[(int64_t)(t + q) * 256] : 0u;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (t + q < nt) hp[(int64_t)(t + q) * 256] = run;
      run += v[q];
    }
  }
  int tot;
  const int base = wg_excl_sum<256, int>((int)run, tmp, &tot);
  hp[0] = (uint32_t)(g.start + base);
}

__global__ void __launch_bounds__(256) k_lscan(Dev d) {
  const int cur = ctr_get(d, CTR_ROUND) & 1;
  for (ItemLoop it(ctr_get(d, CTR_NLARGE)); it.i < it.end; it.i += it.step) lscan_item(d, cur, it.i);
}

// One LSD pass over a large-group tile: stable rank of every row by digit,
// rows placed in LDS in tile-sorted order, then written out so that rows of
// one digit go to consecutive global addresses (coalesced runs).  NT threads
// (NT/64 waves), 4096/NT rows each; the shape is an autotuned choice and
// every shape gives the same result.
template <int NT, bool GEN>
__device__ inline void lscatter_item(const Dev& d, int cur, int shift, int src, int t) {
  const int tid_ = loop_tid();
  constexpr int NW = NT / 64, PER = kTile / NT;
  __shared__ uint32_t cnt[NW][256];
  __shared__ uint32_t lstart[256];   // tile-local start of each digit
  __shared__ uint32_t gstart[256];   // global destination of each digit's first row
  __shared__ uint32_t skey[kTile];
  __shared__ uint32_t sval[kTile];
  __shared__ int32_t tmp[NW + 1];
  const int tid = tid_, wv = tid >> 6, lane = (tid_ & 63);
  const Seg g = seg_get(&d.seg[cur][u32(d.ltile_seg[t])]);
  const int pos = u32(d.ltile_pos[t]);
  const uint64_t off = u64(d.desc[g.blk].off);
  const uint32_t* Ks = src ? d.KEY2 : d.KEY;
  const uint32_t* Ss = src ? d.SA2 : d.SA;
  uint32_t* Kd = src ? d.KEY : d.KEY2;
  uint32_t* Sd = src ? d.SA : d.SA2;
  const int n = min(kTile, g.len - pos);
  const uint64_t base = off + g.start + pos;
  for (int q = tid; q < NW * 256; q += NT) (&cnt[0][0])[q] = 0;
  __syncthreads();
  uint32_t key[PER], val[PER], rk[PER];
  if (GEN) {
    // Round 0's first pass (groups are whole blocks, rows in text order):
    // row i's key is the rotation's first 4 bytes, big-endian, and its SA
    // entry is i with the BWT byte T[(i-1) mod m] on top (as k_bwt_init
    // writes them for the small blocks).  The tile's text, from 4 bytes
    // before it, is staged as words in the (not yet used) skey area.
    const uint8_t* T = d.T + off;
    const int m = g.len;
    uint32_t* tw = skey;
    constexpr int kW = kTile / 4 + 2, kWI = (kW + NT - 1) / NT;
    if (pos >= 4 && pos + 4 * kW - 4 <= m) {
      // inside the block: aligned words (off and pos are multiples of 256),
      // all loads in flight at once
      const uint32_t* T32 = reinterpret_cast<const uint32_t*>(T + pos - 4);
      uint32_t v[kWI];
#pragma unroll
      for (int j = 0; j < kWI; j++) v[j] = tid + NT * j < kW ? T32[tid + NT * j] : 0u;
#pragma unroll
      for (int j = 0; j < kWI; j++)
        if (tid + NT * j < kW) tw[tid + NT * j] = v[j];
    } else {
      // the block's first and last tiles: byte by byte, positions mod m
      for (int w = tid; w < kW; w += NT) {
        const int a = pos + 4 * w - 4;
        uint32_t v = 0;
        for (int j = 0; j < 4; j++) {
          int i = (a + j) % m;
          i = i < 0 ? i + m : i;
          v |= (uint32_t)T[i] << (8 * j);
        }
        tw[w] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = wv * (64 * PER) + q * 64 + lane;
      const int e4 = e + 4;  // byte index of row e in the staged words
      const uint32_t lo = tw[e4 >> 2], hi = tw[(e4 >> 2) + 1];
      const uint32_t x = __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)(e4 & 3));
      const uint32_t pv = (tw[(e4 - 1) >> 2] >> (8 * ((e4 - 1) & 3))) & 255u;
      key[q] = e < n ? __builtin_bswap32(x) : 0u;
      val[q] = e < n ? (uint32_t)(pos + e) | (pv << 24) : 0u;
    }
    __syncthreads();  // (the staged words are overwritten below)
  } else {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = wv * (64 * PER) + q * 64 + lane;
      key[q] = e < n ? Ks[base + e] : 0u;
      val[q] = e < n ? Ss[base + e] : 0u;
    }
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = wv * (64 * PER) + q * 64 + lane;
    rk[q] = wave_rank_digit((key[q] >> shift) & 255u, e < n, cnt[wv]);
  }
  __syncthreads();
  {
    uint32_t c[NW];
    int sum = 0;
    if (tid < 256) {
#pragma unroll
      for (int w = 0; w < NW; w++) { c[w] = cnt[w][tid]; sum += (int)c[w]; }
    }
    int tot;
    const int ls = wg_excl_sum<NT, int>(sum, tmp, &tot);
    if (tid < 256) {
      lstart[tid] = (uint32_t)ls;
      // a tile's prefix within its group plus the group's digit start,
      // which the group's first tile entry holds (k_lscan)
      const int t0 = t - pos / kTile;
      const uint32_t hv = d.hist[(int64_t)t * 256 + tid];
      gstart[tid] = t0 == t ? hv : hv + d.hist[(int64_t)t0 * 256 + tid];
      uint32_t r = (uint32_t)ls;
#pragma unroll
      for (int w = 0; w < NW; w++) { cnt[w][tid] = r; r += c[w]; }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = wv * (64 * PER) + q * 64 + lane;
    if (e < n) {
      const uint32_t p = cnt[wv][(key[q] >> shift) & 255u] + rk[q];
      skey[p] = key[q];
      sval[p] = val[q];
    }
  }
  __syncthreads();
  for (int p = tid; p < n; p += NT) {
    const uint32_t k = skey[p];
    const uint32_t dg = (k >> shift) & 255u;
    const uint32_t dst = gstart[dg] + ((uint32_t)p - lstart[dg]);
    Kd[off + dst] = k;
    Sd[off + dst] = sval[p];
  }
}

template <int NT, bool GEN>
__global__ void __launch_bounds__(NT) k_lscatter(Dev d, int shift, int src) {
  const int cur = ctr_get(d, CTR_ROUND) & 1;
  for (ItemLoop it(ctr_get(d, CTR_LTILES)); it.i < it.end; it.i += it.step) {
    __syncthreads();  // the previous tile's LDS is no longer read
    lscatter_item<NT, GEN>(d, cur, shift, src, it.i);
  }
}

// first / last group boundary (seg-relative) of each large tile
__device__ inline void lsum_item(const Dev& d, int cur, int fin, int t) {
  const int tid_ = loop_tid();
  __shared__ int32_t tmp[8];
  const Seg g = seg_get(&d.seg[cur][u32(d.ltile_seg[t])]);
  const int pos = u32(d.ltile_pos[t]);
  const uint64_t base = u64(d.desc[g.blk].off) + g.start;
  const uint32_t* K = fin ? d.KEY2 : d.KEY;
  const int n = min(kTile, g.len - pos);
  int fb = 0x7fffffff, lb = -1;
  // each key is loaded once (all 16 loads in flight); the key before a
  // lane's row comes from the lane below, lane 0 loads it
  uint32_t kv[16], kp[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = (int)tid_ + 256 * q;
    kv[q] = e < n ? K[base + pos + e] : 0u;
    kp[q] = (tid_ & 63) == 0 && e < n && pos + e > 0 ? K[base + pos + e - 1] : 0u;
  }
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = (int)tid_ + 256 * q;
    const uint32_t below = (uint32_t)__shfl_up((int)kv[q], 1, 64);
    const uint32_t prev = (tid_ & 63) == 0 ? kp[q] : below;
    const int p = pos + e;
    if (e < n && (p == 0 || kv[q] != prev)) { fb = min(fb, p); lb = max(lb, p); }
  }
  fb = wave_min(fb);
  lb = wave_max(lb);
  if ((tid_ & 63) == 0) { tmp[tid_ >> 6] = fb; tmp[4 + (tid_ >> 6)] = lb; }
  __syncthreads();
  if (tid_ == 0) {
    int f = tmp[0], l2 = tmp[4];
    for (int w = 1; w < 4; w++) { f = min(f, tmp[w]); l2 = max(l2, tmp[4 + w]); }
    d.lsum[2 * t] = f;
    d.lsum[2 * t + 1] = l2;
    // Round 0: a whole tile without a group start is a group of more than
    // kTile rows, which round 1 sorts by ranks (the tile path), so the block
    // keeps its ranks from round 0 on (the heads below write them); shorter
    // long groups are found when round 1 is planned.
    if (f == 0x7fffffff && n == kTile && ctr_get(d, CTR_ROUND) == 0 && atomicExch(&d.st[g.blk].rk, 2) != 2)
      atomicAdd(&d.ctr[CTR_KEPT], 1);
  }
}

__global__ void __launch_bounds__(256) k_lsum(Dev d, int fin) {
  const int cur = ctr_get(d, CTR_ROUND) & 1;
  for (ItemLoop it(ctr_get(d, CTR_LTILES)); it.i < it.end; it.i += it.step) {
    __syncthreads();
    lsum_item(d, cur, fin, it.i);
  }
}

// A workgroup of NT threads per tile; thread t owns rows [RPT t, RPT t + RPT).
template <int NT>
__device__ inline void lheads_item(const Dev& d, int cur, int fin, int t) {
  constexpr int RPT = kTile / NT;     // rows per thread
  constexpr int KL = kTile / NT + 1;  // key loads per thread (rows -1 .. kTile)
  constexpr int NW = NT / 64;
  const int tid_ = loop_tid();
  __shared__ int32_t tmp[NW + 1];
  __shared__ int32_t s_carry, s_base;
  __shared__ int32_t xs[NT];
  __shared__ uint32_t sk[kTile + 2 + (kTile + 2) / RPT + 1];
  __shared__ uint32_t ss[kTile + kTile / RPT];
  // per-bucket counters reuse the key staging area once the keys are read
  int32_t* bcnt = reinterpret_cast<int32_t*>(sk);
  static_assert((kStageMax + kRkBucket - 1) / kRkBucket <= kTile, "bucket counters fit the key area");
  const int tid = tid_;
  const int si = u32(d.ltile_seg[t]);
  const Seg g = seg_get(&d.seg[cur][si]);
  const int pos = u32(d.ltile_pos[t]);
  const int kt = pos / kTile;           // tile index within the group
  const int t0 = t - kt;
  const uint64_t off = u64(d.desc[g.blk].off);
  const uint64_t base = off + g.start;
  const uint32_t* K = fin ? d.KEY2 : d.KEY;
  const uint32_t* S = fin ? d.SA2 : d.SA;
  const int n = min(kTile, g.len - pos);
  const int nxt = cur ^ 1;
  // SA values are read only to move them (fin) or for the rank buckets
  const bool kept = u32(d.st[g.blk].rk) == 2;
  const bool need_sa = fin || kept;
  // the tile's keys (with the row before and after) and SA values, loaded
  // before the carry's loads so that both latencies overlap
  uint32_t kv[KL], sv[RPT];
#pragma unroll
  for (int q = 0; q < KL; q++) {
    const int e = tid - 1 + NT * q;
    const int p = pos + e;
    kv[q] = (e <= n && p >= 0 && p < g.len) ? K[base + p] : 0u;
  }
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int e = tid + NT * q;
    sv[q] = need_sa && e < n ? S[base + pos + e] : 0u;
  }
  // carry: last boundary before this tile (largest lb of earlier tiles)
  {
    int lbv = -1;
    for (int j = kt - 1 - tid; j >= 0; j -= NT) {
      int v = d.lsum[2 * (t0 + j) + 1];
      if (v >= 0) { lbv = v; break; }
    }
    lbv = wave_max(lbv);
    if ((tid_ & 63) == 0) tmp[tid >> 6] = lbv;
    __syncthreads();
    if (tid == 0) {
      int c = tmp[0];
      for (int w = 1; w < NW; w++) c = max(c, tmp[w]);
      s_carry = c;
    }
    __syncthreads();
  }
  // stage the keys and SA values in LDS; row e sits at e + e/RPT (one word
  // of skew per RPT rows), so the RPT-row runs the threads own below are
  // read without bank conflicts
  {
#pragma unroll
    for (int q = 0; q < KL; q++) {
      const int e = tid - 1 + NT * q;
      if (e <= n) {
        const int ix = e + 1;
        sk[ix + ix / RPT] = kv[q];
      }
    }
    if (need_sa) {
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int e = tid + NT * q;
        if (e < n) {
          ss[e + e / RPT] = sv[q];
          if (fin) d.SA[base + pos + e] = sv[q];
        }
      }
    }
  }
  __syncthreads();
  const int e0 = tid * RPT;
  uint32_t kk[RPT + 1];
  uint32_t bm = 0;
  int lastB = -1;
  {
    const int ix = e0;  // row e0 - 1
    uint32_t prev = sk[ix + ix / RPT];
#pragma unroll
    for (int q = 0; q <= RPT; q++) {
      const int jx = e0 + q + 1;
      kk[q] = sk[jx + jx / RPT];
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int e = e0 + q;
      const int p = pos + e;
      if (e < n && (p == 0 || kk[q] != prev)) { bm |= 1u << q; lastB = p; }
      prev = kk[q];
    }
  }
  int incl = wg_incl_max<NT, int>(lastB, tmp, s_carry);
  xs[tid] = incl;
  __syncthreads();
  int run = tid > 0 ? xs[tid - 1] : s_carry;
  run = max(run, s_carry);
  int head_of[RPT];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    if (bm & (1u << q)) run = pos + e0 + q;
    head_of[q] = run;
  }
  // Rank updates RK[start] = head go to the block's buckets of kRkBucket
  // starts (k_rk_apply writes them), so that a bucket's updates land in one
  // 32 KiB window that is staged in LDS; only blocks that keep ranks (none
  // in round 0, whose rows are ranked later only if their block needs it).
  if (kept) {
    const int tile_c0 = u32(d.desc[g.blk].tile_c0);
    const int nbk = (u32(d.st[g.blk].m) + kRkBucket - 1) / kRkBucket;
    for (int j = tid; j < nbk; j += NT) bcnt[j] = 0;
    __syncthreads();
    uint32_t lr[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int e = e0 + q;
      lr[q] = e < n ? (uint32_t)atomicAdd(&bcnt[(ss[e + e / RPT] & kSAMask) / kRkBucket], 1) : 0u;
    }
    __syncthreads();
    for (int j = tid; j < nbk; j += NT) {
      const int c = bcnt[j];
      if (c) bcnt[j] = atomicAdd(&d.rkb_cnt[tile_c0 + j * (kRkBucket / kTile)], c);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int e = e0 + q;
      if (e < n) {
        const uint32_t v = ss[e + e / RPT] & kSAMask;
        const int bk = (int)(v / kRkBucket);
        d.RKB[off + (uint64_t)bk * kRkBucket + bcnt[bk] + lr[q]] =
            ((uint64_t)(uint32_t)(g.start + head_of[q]) << 32) | v;
      }
    }
  }
  // groups of >= 2 rows ending in this thread's rows
  uint32_t endm = 0;
  int nout = 0;
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int e = e0 + q;
    const int p = pos + e;
    if (e < n && ((p == g.len - 1) || (kk[q + 1] != kk[q])) && p > head_of[q]) {
      endm |= 1u << q;
      nout++;
    }
  }
  int tot;
  const int ex = wg_excl_sum<NT, int>(nout, tmp, &tot);
  if (tid == 0 && tot > 0) s_base = atomicAdd(&d.ctr[nxt], tot);
  __syncthreads();
  if (endm) {
    int idx = s_base + ex;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      if ((endm >> q) & 1u) {
        if (idx < d.seg_cap) d.seg[nxt][idx] = Seg{g.blk, g.start + head_of[q], pos + e0 + q - head_of[q] + 1, 0};
        else d.ctr[CTR_ERR] = 1;
        idx++;
      }
    }
  }
}

constexpr int kHeadsNT = 512;  // threads per group-head workgroup
// (at most 80 VGPRs: three 512-thread workgroups per CU instead of two; the
// pass is latency-bound, and the third workgroup took it from 8.4 to 7.1 ms
// per batched step)
template <int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) k_lheads(Dev d, int fin) {
  const int cur = ctr_get(d, CTR_ROUND) & 1;
  for (ItemLoop it(ctr_get(d, CTR_LTILES)); it.i < it.end; it.i += it.step) {
    __syncthreads();  // the previous tile's LDS is no longer read
    lheads_item<NT>(d, cur, fin, it.i);
  }
}

// Rank-update buckets: cleared before a round's large groups, applied after
// (nothing to do in a round without large groups).
__global__ void __launch_bounds__(256) k_rk_zero(Dev d) {
  if (ctr_get(d, CTR_LTILES) == 0) return;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < d.ntile_c; i += gridDim.x * 256) d.rkb_cnt[i] = 0;
}

// A bucket's updates are applied to an LDS copy of its window of RK, then
// the touched entries are written out in order.
// (build: only the buckets of the blocks whose ranks are being built)
__device__ inline void rk_apply_item(const Dev& d, int t, bool build) {
  const int tid_ = loop_tid();
  __shared__ uint32_t win[kRkBucket];
  __shared__ uint8_t touched[kRkBucket];
  const int b = u32(d.tc_blk[t]);
  if (build && u32(d.st[b].rk) != 1) return;
  const int k = t - u32(d.desc[b].tile_c0);
  if (k % (kRkBucket / kTile)) return;
  const int cnt = u32(d.rkb_cnt[t]);
  if (cnt == 0) return;
  const uint64_t boff = u64(d.desc[b].off);
  const uint32_t base = (uint32_t)k * kTile;  // first start of the bucket
  for (int i = tid_; i < kRkBucket; i += 256) touched[i] = 0;
  __syncthreads();
  const uint64_t* src = d.RKB + boff + base;
  for (int i0 = tid_; i0 < cnt; i0 += 4096) {
    uint64_t p[16];  // 16 loads in flight per thread
#pragma unroll
    for (int q = 0; q < 16; q++) p[q] = i0 + 256 * q < cnt ? src[i0 + 256 * q] : ~0ull;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (p[q] != ~0ull) {
        const uint32_t o = (uint32_t)p[q] - base;
        win[o] = (uint32_t)(p[q] >> 32);
        touched[o] = 1;
      }
    }
  }
  __syncthreads();
  const int span = min(kRkBucket, u32(d.st[b].m) - (int)base);
  for (int i = tid_; i < span; i += 256)
    if (touched[i]) d.RK[boff + base + i] = win[i];
}

// build = 0: the round's large-group heads (none to apply while no block
// keeps ranks); 1: blocks whose ranks are built
__global__ void __launch_bounds__(256) k_rk_apply(Dev d, int build) {
  if (build ? ctr_get(d, CTR_RKN) == 0 : (ctr_get(d, CTR_LTILES) == 0 || ctr_get(d, CTR_KEPT) == 0)) return;
  // a bucket covers kRkBucket / kTile capacity tiles; its first tile applies it
  for (ItemLoop it(build ? rk_items(d) : d.ntile_c); it.i < it.end; it.i += it.step) {
    __syncthreads();
    const int t = build ? rk_item_tile(d, it.i) : it.i;
    if (t >= 0) rk_apply_item(d, t, build != 0);
  }
}

// ---------------------------------------------------------- rank building
// Round 2: a pack with a group of more than kCountMax rows is sorted by
// composite rank keys, so every block with a group in it needs its ranks
// (from round 3 on, k_mk_large flags every listed block).  Round 1 sorts
// such packs by text keys, in a 64-bit sort about three times as costly per
// row as building a block's ranks per block row, and a block with many rows
// there most likely needs its ranks in a later round: here round 1 counts
// those rows per block, and k_rk_pick flags the blocks with more than
// m / kRkShare of them.
__global__ void __launch_bounds__(256) k_pack_flags(Dev d) {
  __shared__ int32_t smax[4];
  const int round = ctr_get(d, CTR_ROUND), cur = round & 1;
  if (round != 1 && round != 2) return;
  const int tid = loop_tid();
  for (ItemLoop it(round_packs(d)); it.i < it.end; it.i += it.step) {
    const int i0 = u32(d.packs[it.i]), i1 = u32(d.packs[it.i + 1]);
    int mx = 0;
    for (int i = i0 + tid; i < i1; i += 256) {
      const int len = d.seg[cur][i].len;
      if (len <= kHalf) mx = max(mx, len);
    }
    mx = wave_max(mx);
    __syncthreads();  // the previous item's maxima are read
    if ((tid & 63) == 0) smax[tid >> 6] = mx;
    __syncthreads();
    if (max(max(smax[0], smax[1]), max(smax[2], smax[3])) > kCountMax) {
      // (the key gather of rounds 1-2 and round 1's second sort launch take
      // only the listed packs)
      if (tid == 0) d.rpk[atomicAdd(&d.ctr[CTR_RPK], 1)] = it.i;
      for (int base = i0; base < i1; base += 256) {  // whole waves (the sums)
        const int i = base + tid;
        Seg g{0, 0, 0, 0};
        if (i < i1) g = d.seg[cur][i];
        const int len = g.len <= kHalf ? g.len : 0;
        if (round == 2) {
          if (len > 0) rk_need(d, g.blk);
          continue;
        }
        // a wave's groups mostly share a block: one add for the wave then
        const int b0 = __shfl(g.blk, 0, 64);
        if (__ballot(len > 0 && g.blk != b0) == 0) {
          const int sum = wave_sum(len);
          if ((tid & 63) == 0 && sum > 0) atomicAdd(&d.st[b0].rk_rows, sum);
        } else if (len > 0) {
          atomicAdd(&d.st[g.blk].rk_rows, len);
        }
      }
    }
  }
}

// Round 1: blocks with many rows in packs sorted by composite keys build
// their ranks (k_pack_flags).
__global__ void __launch_bounds__(256) k_rk_pick(Dev d) {
  if (ctr_get(d, CTR_ROUND) != 1) return;
  for (int b = blockIdx.x * 256 + threadIdx.x; b < d.nblk; b += gridDim.x * 256)
    if ((int64_t)d.st[b].rk_rows * kRkShare > d.st[b].m) rk_need(d, b);
}

// Bucket counters of the blocks whose ranks are built.
__global__ void __launch_bounds__(256) k_rkb_zero(Dev d) {
  const int n = ctr_get(d, CTR_RKN);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n * d.tiles_max; i += gridDim.x * 256) {
    const int b = d.rk_list[i / d.tiles_max], k = i % d.tiles_max;
    if (k < tiles_for(d.desc[b].cap)) d.rkb_cnt[d.desc[b].tile_c0 + k] = 0;
  }
}

// A block's ranks are built in three steps: every row's rank in a row-order
// array H (KEY2, free between rounds) is the row itself (k_rk_self), then
// the group start for the rows of the round's groups (k_rk_groups, writes
// contiguous per group), and RK[SA[i]] = H[i] goes through the rank buckets
// as the group heads' updates do (k_rk_rows, then k_rk_apply).
__global__ void __launch_bounds__(256) k_rk_self(Dev d) {
  if (ctr_get(d, CTR_RKN) == 0) return;
  for (ItemLoop it(rk_items(d)); it.i < it.end; it.i += it.step) {
    const int t = rk_item_tile(d, it.i);
    if (t < 0) continue;
    const int b = u32(d.tc_blk[t]);
    const uint64_t off = u64(d.desc[b].off);
    const int m = u32(d.st[b].m);
    const int start = (t - u32(d.desc[b].tile_c0)) * kTile;
    for (int e = (int)threadIdx.x; e < kTile && start + e < m; e += 256) d.KEY2[off + start + e] = (uint32_t)(start + e);
  }
}

__device__ inline void rk_rows_item(const Dev& d, int t) {
  const int tid = loop_tid();
  __shared__ int32_t bcnt[(kStageMax + kRkBucket - 1) / kRkBucket];
  const int b = u32(d.tc_blk[t]);
  if (u32(d.st[b].rk) != 1) return;
  const uint64_t off = u64(d.desc[b].off);
  const int tile_c0 = u32(d.desc[b].tile_c0);
  const int m = u32(d.st[b].m);
  const int start = (t - tile_c0) * kTile;
  if (start >= m) return;
  const int n = min(kTile, m - start);
  const int nbk = (m + kRkBucket - 1) / kRkBucket;
  for (int j = tid; j < nbk; j += 256) bcnt[j] = 0;
  uint32_t sv[kTile / 256], hv[kTile / 256];
#pragma unroll
  for (int q = 0; q < kTile / 256; q++) {
    const int e = tid + 256 * q;
    sv[q] = e < n ? d.SA[off + start + e] & kSAMask : 0u;
    hv[q] = e < n ? d.KEY2[off + start + e] : 0u;
  }
  __syncthreads();
  uint32_t lr[kTile / 256];
#pragma unroll
  for (int q = 0; q < kTile / 256; q++)
    lr[q] = tid + 256 * q < n ? (uint32_t)atomicAdd(&bcnt[sv[q] / kRkBucket], 1) : 0u;
  __syncthreads();
  for (int j = tid; j < nbk; j += 256) {
    const int c = bcnt[j];
    if (c) bcnt[j] = atomicAdd(&d.rkb_cnt[tile_c0 + j * (kRkBucket / kTile)], c);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kTile / 256; q++) {
    const int e = tid + 256 * q;
    if (e < n) {
      const int bk = (int)(sv[q] / kRkBucket);
      d.RKB[off + (uint64_t)bk * kRkBucket + bcnt[bk] + lr[q]] = ((uint64_t)hv[q] << 32) | sv[q];
    }
  }
}

__global__ void __launch_bounds__(256) k_rk_rows(Dev d) {
  if (ctr_get(d, CTR_RKN) == 0) return;
  for (ItemLoop it(rk_items(d)); it.i < it.end; it.i += it.step) {
    __syncthreads();  // the previous tile's counters are no longer read
    const int t = rk_item_tile(d, it.i);
    if (t >= 0) rk_rows_item(d, t);
  }
}

// The rows of the round's groups in blocks being built take their group's
// start in H (small groups one per thread, large groups by tile).
__global__ void __launch_bounds__(256) k_rk_groups(Dev d) {
  if (ctr_get(d, CTR_RKN) == 0) return;
  const int cur = ctr_get(d, CTR_ROUND) & 1, nseg = ctr_get(d, CTR_NSEG), ntl = ctr_get(d, CTR_LTILES);
  const int nch = (nseg + 255) / 256;
  for (ItemLoop it(nch + ntl); it.i < it.end; it.i += it.step) {
    if (it.i < nch) {
      const int i = it.i * 256 + (int)threadIdx.x;
      if (i < nseg) {
        const Seg g = d.seg[cur][i];
        if (g.len <= kHalf && d.st[g.blk].rk == 1) {
          const uint64_t off = d.desc[g.blk].off;
          for (int j = 0; j < g.len; j++) d.KEY2[off + g.start + j] = (uint32_t)g.start;
        }
      }
    } else {
      const int t = it.i - nch;
      const Seg g = seg_get(&d.seg[cur][u32(d.ltile_seg[t])]);
      if (u32(d.st[g.blk].rk) != 1) continue;
      const int pos = u32(d.ltile_pos[t]);
      const uint64_t off = u64(d.desc[g.blk].off);
      const int n = min(kTile, g.len - pos);
      for (int e = (int)threadIdx.x; e < n; e += 256) d.KEY2[off + g.start + pos + e] = (uint32_t)g.start;
    }
  }
}

// then those blocks keep their ranks
__global__ void __launch_bounds__(256) k_rk_kept(Dev d) {
  const int n = ctr_get(d, CTR_RKN);
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) d.st[d.rk_list[j]].rk = 2;
  if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0) atomicAdd(&d.ctr[CTR_KEPT], n);
}

void launch_bwt_init(const Dev& d, hipStream_t s) {
  k_bwt_init<<<d.ntile_c, 256, 0, s>>>(d);
}

namespace {
// persistent grid of a kernel: every CU filled to the kernel's residency,
// a multiple of the 8 XCDs
template <typename K>
int grid_for(K kernel, int threads, int cus) {
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
  return std::max(8, (cus * per + 7) / 8 * 8);
}
}  // namespace

SortGrids sort_grids(int device) {
  SortGrids g;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
  g.light = grid_for(k_lhist<false>, 256, cus);
  g.keys = grid_for(k_keys, 256, cus);
  g.small = grid_for(k_sort_small, 256, cus);
  g.small64 = grid_for(k_sort_small64, 256, cus);
  g.scatter[0] = grid_for(k_lscatter<256, false>, 256, cus);
  g.scatter[1] = grid_for(k_lscatter<512, false>, 512, cus);
  g.scatter[2] = grid_for(k_lscatter<1024, false>, 1024, cus);
  g.heads = grid_for(k_lheads<kHeadsNT>, kHeadsNT, cus);
  g.apply = grid_for(k_rk_apply, 256, cus);
  g.apply = std::min(g.apply, grid_for(k_rk_rows, 256, cus));
  return g;
}

namespace {
int round8(int n) { return n < 8 ? 8 : (n + 7) / 8 * 8; }
}  // namespace

void launch_bwt_round(const Dev& d, const SortGrids& g, bool first, int late_bits, int scatter_nt,
                      hipStream_t s, const hipEvent_t* pass_ev, const int32_t* exact) {
  // Without `exact`, the round plans itself.  Round 0's work is bounded by
  // the plan (at most one pack, list entry and large group per block, at
  // most one tile per capacity tile), so its kernels get one workgroup per
  // possible item and the ones past the device-side count exit at once;
  // later rounds run on persistent grids.  With `exact` (the round's
  // counters, read back after launch_bwt_plan), each kernel gets one
  // workgroup per work item and is skipped when it has none.
  int npacks = 0, ntiles = 0, nlarge = 0, nseg = 0;
  if (exact) {
    npacks = (exact[CTR_SMALL] + kHalf - 1) / kHalf;
    ntiles = exact[CTR_LTILES];
    nlarge = exact[CTR_NLARGE];
    nseg = exact[CTR_NSEG];
  } else {
    launch_bwt_plan(d, s);
    if (first) {
      npacks = nlarge = nseg = d.nblk;
      ntiles = d.ntile_c;
    }
  }
  const bool sized = exact || first;
  auto grid = [&](int persistent, int items) { return sized ? round8(items) : persistent; };
  k_mk_packs<<<sized ? (npacks + 1 + 255) / 256 : g.light, 256, 0, s>>>(d);
  k_mk_large<<<sized ? std::max(1, (nseg + 255) / 256) : g.light, 256, 0, s>>>(d);
  if (!first) {
    // ranks of the blocks that need them from this round on (the kernels
    // return at once when no block is flagged)
    k_pack_flags<<<g.keys, 256, 0, s>>>(d);
    k_rk_pick<<<std::max(1, std::min(g.light, (d.nblk + 255) / 256)), 256, 0, s>>>(d);
    k_rkb_zero<<<g.light, 256, 0, s>>>(d);
    k_rk_self<<<g.keys, 256, 0, s>>>(d);
    k_rk_groups<<<g.keys, 256, 0, s>>>(d);
    k_rk_rows<<<g.apply, 256, 0, s>>>(d);
    k_rk_apply<<<g.apply, 256, 0, s>>>(d, 1);
    k_rk_kept<<<std::max(1, std::min(g.light, (d.nblk + 255) / 256)), 256, 0, s>>>(d);
  }
  if (!first && (!exact || npacks + ntiles > 0)) {  // round-0 keys come from the text
    probe_begin("k_keys", s);
    k_keys<<<grid(g.keys, npacks + ntiles), 256, 0, s>>>(d);
    probe_end("k_keys", s);
  }
  // composite key of a pack sort: group index in the pack (< 1024) above
  // the key bits
  const int key_bits = first ? 32 : late_bits;
  if (!exact || npacks > 0) {
    probe_begin("k_sort_small", s);
    if (key_bits + 10 <= 32) k_sort_small<<<grid(g.small, npacks), 256, 0, s>>>(d, key_bits);
    else k_sort_small64<<<grid(g.small64, npacks), 256, 0, s>>>(d, key_bits, 0);
    // round 1's packs sorted by composite text keys (returns in other rounds)
    if (!first) k_sort_small64<<<grid(g.small64, npacks), 256, 0, s>>>(d, 32, 1);
    probe_end("k_sort_small", s);
  }
  if (exact && ntiles == 0) return;
  const int sv = scatter_nt == 256 ? 0 : (scatter_nt == 1024 ? 2 : 1);
  const int sg = grid(g.scatter[sv], ntiles);
  const int npass = (key_bits + 7) / 8;
  for (int p = 0; p < npass; p++) {
    if (pass_ev) (void)hipEventRecord(pass_ev[p], s);
    // round 0's first pass reads the text (k_bwt_init leaves KEY / SA of the
    // large blocks unwritten); it is not one of the probed passes
    const bool gen = first && p == 0;
    if (gen) k_lhist<true><<<grid(g.light, ntiles), 256, 0, s>>>(d, 8 * p, p & 1);
    else k_lhist<false><<<grid(g.light, ntiles), 256, 0, s>>>(d, 8 * p, p & 1);
    k_lscan<<<grid(g.light, nlarge), 256, 0, s>>>(d);
    if (gen) {
      if (sv == 0) k_lscatter<256, true><<<sg, 256, 0, s>>>(d, 8 * p, p & 1);
      else if (sv == 2) k_lscatter<1024, true><<<sg, 1024, 0, s>>>(d, 8 * p, p & 1);
      else k_lscatter<512, true><<<sg, 512, 0, s>>>(d, 8 * p, p & 1);
      continue;
    }
    probe_begin("k_lscatter", s);
    if (sv == 0) k_lscatter<256, false><<<sg, 256, 0, s>>>(d, 8 * p, p & 1);
    else if (sv == 2) k_lscatter<1024, false><<<sg, 1024, 0, s>>>(d, 8 * p, p & 1);
    else k_lscatter<512, false><<<sg, 512, 0, s>>>(d, 8 * p, p & 1);
    probe_end("k_lscatter", s);
  }
  if (pass_ev) (void)hipEventRecord(pass_ev[npass], s);
  const int fin = npass & 1;
  k_rk_zero<<<sized ? std::max(1, (d.ntile_c + 255) / 256) : g.light, 256, 0, s>>>(d);
  k_lsum<<<grid(g.light, ntiles), 256, 0, s>>>(d, fin);
  probe_begin("k_lheads", s);
  k_lheads<kHeadsNT><<<grid(g.heads, ntiles), kHeadsNT, 0, s>>>(d, fin);
  probe_end("k_lheads", s);
  k_rk_apply<<<grid(g.apply, d.ntile_c), 256, 0, s>>>(d, 0);
}

}  // namespace bzg
