// huff.hip - Huffman table build, selectors, block headers, payload packing
// and .bz2 stream assembly for a whole batch.
//
// Table build (reference: src/compression/huffman.cu:148-364, which follows
// libbz2's sendMTFValues): n_groups from s, an initial split of the alphabet
// into n_groups frequency ranges (cost 0 inside / 15 outside), then 4
// iterations of: per 50-symbol group pick the first table of strictly
// minimal cost and count the group's symbols for it, then rebuild every
// table's code lengths with the libbz2 heap method (max length 17).
// Codes are canonical by (length, symbol).
//
// Bit output: every writer owns a contiguous bit range of a zeroed word
// array; words fully inside the range are plain stores, the (at most two)
// words shared with a neighbour are atomicOr.  Words are MSB-first u32
// (src/compression/huffman_encode.cu:30-71); the final stream stores them
// byte-swapped so memory holds the .bz2 bytes in order.
#include "pipeline.h"

namespace bzg {

// ---------------------------------------------------------------- bit out
struct BitOut {
  uint32_t* w;       // word array
  int64_t pos;       // next bit position
  int64_t lo, hi;    // owned bit range [lo, hi)
  uint64_t acc;      // pending bits, right-aligned
  int nacc;
  bool swap;         // store byte-swapped (stream output)

  __device__ void init(uint32_t* words, int64_t start, int64_t end, bool bswap) {
    w = words; pos = start; lo = start; hi = end; acc = 0; nacc = 0; swap = bswap;
  }
  __device__ void store(int64_t wi, uint32_t v, bool full) {
    if (swap) v = __builtin_bswap32(v);
    if (full) w[wi] = v;
    else if (v) atomicOr(&w[wi], v);
  }
  __device__ void put(uint32_t v, int n) {  // 0 <= n <= 32
    if (n == 0) return;
    acc = (acc << n) | (uint64_t)(n == 32 ? v : (v & ((1u << n) - 1)));
    nacc += n;
    int room = 32 - (int)(pos & 31);
    while (nacc >= room) {
      uint32_t bits = (uint32_t)(acc >> (nacc - room));
      if (room < 32) bits &= (1u << room) - 1;
      int64_t wi = pos >> 5;
      store(wi, bits, wi * 32 >= lo && wi * 32 + 32 <= hi);
      pos += room;
      nacc -= room;
      acc = nacc ? (acc & ((1ull << nacc) - 1)) : 0ull;
      room = 32;
    }
  }
  __device__ void finish() {  // trailing partial word (shared with the right)
    if (nacc > 0) {
      int sh = (int)(pos & 31);
      store(pos >> 5, (uint32_t)acc << (32 - sh - nacc), false);
      pos += nacc;
      nacc = 0;
      acc = 0;
    }
  }
};

// ------------------------------------------------------------ table build
__global__ void __launch_bounds__(64) k_huff_init(Dev d) {
  __shared__ int32_t fq[264];
  __shared__ int32_t pgs[6], pge[6];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  BlockState* st = d.st + b;
  const int s = st->s;
  const int alpha = st->n_in_use + 2;
  const int ng = s < 200 ? 2 : s < 600 ? 3 : s < 1200 ? 4 : s < 2400 ? 5 : 6;
  const int32_t* freq = d.freq + (int64_t)b * 260;
  {
    int f[5];  // the counts staged in LDS, all loads in flight
#pragma unroll
    for (int r = 0; r < 5; r++) f[r] = tid + 64 * r < 260 ? freq[tid + 64 * r] : 0;
#pragma unroll
    for (int r = 0; r < 5; r++)
      if (tid + 64 * r < 264) fq[tid + 64 * r] = f[r];
  }
  // zero this block's header words (header writers use atomicOr at edges)
  uint32_t* hw = d.hdr + (int64_t)b * d.hdr_stride;
  for (int64_t q = tid; q < d.hdr_stride; q += 64) hw[q] = 0;
  __syncthreads();
  if (tid == 0) {
    st->n_groups = ng;
    st->n_sel = (s + kGSize - 1) / kGSize;
    // initial split, huffman.cu:148-185 (bound `ge < alphabet_size`, :160);
    // table part-1 covers symbols [gs, ge]
    int gs = 0, rem = s;
    for (int part = ng; part > 0; part--) {
      int tf = rem / part, ge = gs - 1, a = 0;
      while (a < tf && ge < alpha) { ge++; a += fq[ge]; }
      if (ge > gs && part != ng && part != 1 && ((ng - part) % 2 == 1)) { a -= fq[ge]; ge--; }
      pgs[part - 1] = gs;
      pge[part - 1] = ge;
      gs = ge + 1;
      rem -= a;
    }
  }
  __syncthreads();
  uint8_t* len = d.len + (int64_t)b * 6 * 260;
  for (int q = tid; q < 6 * 260; q += 64) {
    const int tb = q / 260, v = q - 260 * tb;
    len[q] = (tb < ng && v < alpha && v >= pgs[tb] && v <= pge[tb]) ? 0 : 15;
  }
}

// libbz2 code-length method (huffman.cu:51-124): binary min-heap on
// weight = freq<<8 | depth with the same sift rules, so ties resolve exactly
// as in the reference.  One wave per (block, table); the heap lives in LDS
// and every heap operation is resolved a window of levels at a time: the
// lanes load the window, and sibling minima, path membership and the
// "tmp < entry" stop test become ballots, so one LDS round trip covers five
// levels of a sift.  Depths come from pointer jumping over the parent array.
namespace {

// Orders one wave's LDS accesses across its lanes (a wave's LDS operations
// execute in order; this keeps the compiler from moving them).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sift-down of `tmp` from the root (DOWNHEAP(1)); heap holds nHeap entries.
// Lane l < 62 covers window level k = log2(l+2), index j = l+2-2^k below zz;
// siblings are lanes 2i, 2i+1.  An entry lies on the min-child path when it
// and all its window ancestors are the chosen child of their pair; the path
// stops above the first path entry that tmp is strictly lighter than.
struct WinLane {
  int k, j;
  uint64_t am;  // window lanes of this lane's ancestors (levels 1..k-1)
};

__device__ inline WinLane window_lane(int lane) {
  WinLane w;
  w.k = 31 - __clz(lane + 2);
  w.j = lane + 2 - (1 << w.k);
  w.am = 0;
  for (int kp = 1; kp < w.k; kp++) w.am |= 1ull << ((1 << kp) - 2 + (w.j >> (w.k - kp)));
  return w;
}

constexpr int kHeapCap = 264;
constexpr uint64_t kWinAct = (1ull << 62) - 1;
constexpr uint64_t kWinOdd = 0xAAAAAAAAAAAAAAAAull & kWinAct;
constexpr uint64_t kWinEven = 0x5555555555555555ull & kWinAct;

__device__ inline int2 readlane2(int2 v, int l) {
  return make_int2(__builtin_amdgcn_readlane(v.x, l), __builtin_amdgcn_readlane(v.y, l));
}
__device__ inline int2 uniform2(int2 v) {
  return make_int2(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y));
}

// Returns the new root; *pre receives the entry at position pre_pos after the
// sift (pre_pos is the heap's last position: a leaf, so the sift changes it
// only by ending there).
__device__ inline int2 heap_down(int2* heap, int nHeap, int2 tmp, int lane, const WinLane& wl, int pre_pos,
                                 int2* pre) {
  const int2 pv = uniform2(heap[pre_pos]);  // read with the first window
  int zz = 1, cur;
  int2 root = tmp;
  bool first = true;
  for (;;) {
    const int p = (zz << wl.k) + wl.j;
    const int2 e = heap[min(p, kHeapCap - 1)];  // entries past nHeap are masked below
    const int sib = __builtin_amdgcn_mov_dpp(e.x, 0xB1, 0xF, 0xF, false);  // lane ^ 1
    const uint64_t E = __ballot(p <= nHeap) & kWinAct;
    // the right entry is taken when it exists and is strictly lighter
    const uint64_t R = __ballot(e.x < sib) & E & kWinOdd;
    const uint64_t W = R | (E & kWinEven & ~(R >> 1));
    const uint64_t P = __ballot((W & wl.am) == wl.am) & W;
    const uint64_t S = P & __ballot(tmp.x < e.x);
    // path entries above the first stop move up; without a stop the sift
    // goes on below the window's last level if the path reaches it
    uint64_t moved = P;
    uint32_t more = (uint32_t)(P >> 30);
    if (S) {
      moved = P & ((S & (~S + 1)) - 1);
      more = 0;
    }
    if (first) {
      if (moved & 3ull) root = readlane2(e, __builtin_ctzll(moved));
      first = false;
    }
    cur = moved ? __builtin_amdgcn_readlane(p, 63 - __builtin_clzll(moved)) : zz;
    // the next window reads only below cur, which these writes never touch
    if (__builtin_amdgcn_inverse_ballot_w64(moved)) heap[(zz << (wl.k - 1)) + (wl.j >> 1)] = e;
    if (!more) break;
    zz = cur;
  }
  if (lane == 0) heap[cur] = tmp;
  wave_lds_sync();
  *pre = cur == pre_pos ? tmp : pv;
  return root;
}

// Sift-up of `tmp` placed at position zz (UPHEAP(zz)): lane q holds the
// ancestor zz >> (q+1); every ancestor heavier than tmp moves one level down.
// With TRACK, returns the new root and *last receives the new entry at zz.
template <bool TRACK>
__device__ inline int2 heap_up(int2* heap, int zz, int2 tmp, int lane, int2* last) {
  uint64_t* h = reinterpret_cast<uint64_t*>(heap);  // whole entries, one load each
  const int a = lane < 9 ? (zz >> (lane + 1)) : 0;
  const uint64_t ev = h[a];
  const int2 e = make_int2((int)(uint32_t)ev, (int)(uint32_t)(ev >> 32));
  const uint64_t M = __ballot(a >= 1 && tmp.x < e.x);
  const int c = __builtin_ctzll(~M);
  if (lane < c) h[zz >> lane] = ev;
  if (lane == 0) heap[zz >> c] = tmp;
  wave_lds_sync();
  if (!TRACK) return tmp;
  const int L = 31 - __clz(zz);  // ancestors a_0 .. a_(L-1); a_(L-1) is the root
  *last = c >= 1 ? readlane2(e, 0) : tmp;
  return c == L ? tmp : readlane2(e, L - 1);
}

// LDS of one table's code-length build (the leaf weights live in the
// table's symbol-count array, converted in place).
struct HeapLds {
  __align__(16) int2 heap[kHeapCap];
  int16_t anc[520];  // pointer-jumping ancestors (node ids <= 2*258-1)
  int16_t dep[520];
};
// code_too_long's two queues reuse heap/anc/dep: kQueueCap entries each
constexpr int kQueueCap = 384;
static_assert(sizeof(HeapLds) >= 2 * kQueueCap * 4, "queue scratch exceeds the heap arrays");

// Whether the tree built from the leaf weights W[0..alpha) (leaf i+1 has
// weight W[i]) is deeper than 17 levels.
// Any min-priority queue pops the same weight values in the same order
// (only which of two equal weights comes first depends on the queue), and
// the depth byte of a merged weight is one more than its deeper input, so the
// tree height is fixed by the values alone: the leaves are ranked by
// (weight, id), and a two-queue merge over the sorted leaves and the merged
// nodes (which come out in non-decreasing order) follows the heap's merges.
// Each queue is read through a 64-entry register window (one entry per
// lane) that moves on every 64 pops; merged nodes also enter the window of
// their queue when it covers them.
__device__ bool code_too_long(HeapLds& H, const int32_t* W, int alpha, int lane) {
  int32_t* S = reinterpret_cast<int32_t*>(&H);  // sorted leaf weights
  int32_t* C = S + kQueueCap;                   // merged weights, in order
  uint32_t key[5];
  int rk[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int i = lane + 64 * q + 1;
    key[q] = i <= alpha ? ((uint32_t)(W[i - 1] >> 8) << 9) | (uint32_t)i : 0xFFFFFFFFu;
    rk[q] = 0;
  }
  // rank of every leaf key among all keys (keys are distinct)
#pragma unroll
  for (int qj = 0; qj < 5; qj++) {
    if (64 * qj >= alpha) break;
    const int nj = min(64, alpha - 64 * qj);
    for (int l = 0; l < nj; l++) {
      const uint32_t kj = (uint32_t)__builtin_amdgcn_readlane((int)key[qj], l);
#pragma unroll
      for (int q = 0; q < 5; q++) rk[q] += kj < key[q] ? 1 : 0;
    }
  }
#pragma unroll
  for (int q = 0; q < 5; q++)
    if (lane + 64 * q + 1 <= alpha) S[rk[q]] = (int32_t)(key[q] >> 9) << 8;
  wave_lds_sync();
  int i1 = 0, i2 = 0, n2 = 0;
  int winS = S[lane], winC = 0;
  int sh = __builtin_amdgcn_readlane(winS, 0), ch = 0;  // queue heads
  bool deep = false;
  for (int step = 1; step < alpha; step++) {
    int ab[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      if (i1 < alpha && (i2 >= n2 || sh <= ch)) {
        ab[t] = sh;
        i1++;
        if ((i1 & 63) == 0) winS = S[i1 + lane];
        sh = __builtin_amdgcn_readlane(winS, i1 & 63);
      } else {
        ab[t] = ch;
        i2++;
        if ((i2 & 63) == 0) winC = C[i2 + lane];
        ch = __builtin_amdgcn_readlane(winC, i2 & 63);
      }
    }
    const int da = ab[0] & 0xff, db = ab[1] & 0xff;
    const int dn = 1 + (da > db ? da : db);
    if (dn > 17) {
      deep = true;
      break;
    }
    const int c = ((ab[0] & 0xffffff00) + (ab[1] & 0xffffff00)) | dn;
    if (lane == 0) C[n2] = c;
    if ((n2 >> 6) == (i2 >> 6) && lane == (n2 & 63)) winC = c;
    if (n2 == i2) ch = c;
    n2++;
  }
  wave_lds_sync();
  return deep;
}

// Code lengths of one table from its symbol counts W[0..alpha) (LDS, turned
// into the leaf weights in place), by one wave; len[0..alpha) (global)
// receives them.  The weight halving rounds of libbz2 (huffman.cu:110-123)
// are decided by code_too_long before the one exact build.
__device__ void build_lengths(HeapLds& H, int32_t* W, uint8_t* len, int alpha, int lane) {
  const WinLane wl = window_lane(lane);
  for (int i = lane; i < alpha; i += 64) {
    const int f = W[i];
    W[i] = (f == 0 ? 1 : f) << 8;
  }
  wave_lds_sync();
  while (code_too_long(H, W, alpha, lane)) {
    for (int i = lane; i < alpha; i += 64) {
      int j = W[i] >> 8;
      j = 1 + (j / 2);
      W[i] = j << 8;
    }
    wave_lds_sync();
  }
  while (true) {
    int nNodes = alpha, nHeap = 0;
    for (int i0 = 1; i0 <= alpha; i0 += 64) {
      const int lw = i0 + lane <= alpha ? W[i0 + lane - 1] : 0;
      const int n = min(64, alpha + 1 - i0);
      for (int q = 0; q < n; q++) {
        nHeap++;
        heap_up<false>(H.heap, nHeap, make_int2(__builtin_amdgcn_readlane(lw, q), i0 + q), lane, nullptr);
      }
    }
    int2 root = uniform2(H.heap[1]), last = uniform2(H.heap[nHeap]);
    while (nHeap > 1) {
      const int2 n1 = root;
      nHeap--;
      int2 t2;
      const int2 n2 = heap_down(H.heap, nHeap, last, lane, wl, nHeap, &t2);
      nHeap--;
      int2 unused;
      heap_down(H.heap, nHeap, t2, lane, wl, 0, &unused);
      nNodes++;
      if (lane == 0) {
        H.anc[n1.y] = (int16_t)nNodes;
        H.anc[n2.y] = (int16_t)nNodes;
      }
      const int w1 = n1.x, w2 = n2.x;
      const int d1 = w1 & 0xff, d2 = w2 & 0xff;
      const int2 nn = make_int2(((w1 & 0xffffff00) + (w2 & 0xffffff00)) | (1 + (d1 > d2 ? d1 : d2)), nNodes);
      nHeap++;
      root = heap_up<true>(H.heap, nHeap, nn, lane, &last);
    }
    // depth of every node by pointer jumping (root: ancestor itself, depth
    // 0), in place: a node's ancestor has a larger id, so it lies in the
    // same 64-node chunk (whose loads all precede its stores) or in a later
    // one, and every step reads the previous step's values, as a second
    // buffer would give
    for (int v = lane + 1; v <= nNodes; v += 64) {
      const bool root = v == nNodes;
      if (root) H.anc[v] = (int16_t)v;
      H.dep[v] = root ? 0 : 1;
    }
    wave_lds_sync();
    for (int it = 0; it < 9; it++) {
      for (int v = lane + 1; v <= nNodes; v += 64) {
        const int a = H.anc[v];
        const int da = H.dep[a], aa = H.anc[a];
        const int dv = H.dep[v];
        __builtin_amdgcn_wave_barrier();
        H.dep[v] = (int16_t)(dv + da);
        H.anc[v] = (int16_t)aa;
      }
      wave_lds_sync();
    }
    bool tooLong = false;
    for (int i = lane + 1; i <= alpha; i += 64) {
      const int j = H.dep[i];
      len[i - 1] = (uint8_t)j;
      if (j > 17) tooLong = true;
    }
    if (!__ballot(tooLong)) break;
    for (int i = lane; i < alpha; i += 64) {
      int j = W[i] >> 8;
      j = 1 + (j / 2);
      W[i] = j << 8;
    }
    wave_lds_sync();
  }
  wave_lds_sync();
}

}  // namespace

// The four selection / code-length iterations of one block in one
// workgroup of six waves: all threads assign 50-symbol groups (costs from
// 64-bit packed lengths, first strict minimum, counts in LDS), then wave t
// rebuilds table t's lengths if the set of groups that chose table t
// changed (else its counts, and so its lengths, are the previous
// iteration's).  An iteration in which no group changed its table is a
// fixed point: the later ones would repeat it, and they are skipped.  A
// block's tables never wait for another block's.
// Launch order of the code-length builds: blocks by descending alphabet
// size (the heap work grows with it).  More blocks than CUs means some CUs
// run two builds at once; the builds launched first are paired with those
// launched last, so the largest alphabets share a CU with the smallest,
// which finish early.  A counting sort in one workgroup; the order within
// an alphabet size does not matter.
__global__ void __launch_bounds__(256) k_huff_order(Dev d) {
  __shared__ int32_t cnt[260];
  const int tid = threadIdx.x;
  for (int q = tid; q < 260; q += 256) cnt[q] = 0;
  __syncthreads();
  for (int b = tid; b < d.nblk; b += 256) atomicAdd(&cnt[259 - min(d.st[b].n_in_use + 2, 259)], 1);
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int q = 0; q < 260; q++) { const int c = cnt[q]; cnt[q] = run; run += c; }
  }
  __syncthreads();
  for (int b = tid; b < d.nblk; b += 256) d.horder[atomicAdd(&cnt[259 - min(d.st[b].n_in_use + 2, 259)], 1)] = b;
}

// 6 waves: all assign groups, wave t builds table t.  31 KB of LDS (the
// current lengths stay in global memory, the counts become the weights in
// place) and at most 64 VGPRs let five blocks share a CU, so that the heap
// chains of different blocks overlap.
constexpr int kHuffThreads = 384;
__global__ void __launch_bounds__(kHuffThreads) __attribute__((amdgpu_waves_per_eu(8))) k_huff_build(Dev d) {
  __shared__ int32_t rf[6][260];
  __shared__ HeapLds H[6];
  __shared__ int32_t s_chg;  // tables whose set of groups changed this iteration
  // the six lengths of a symbol packed in one word, used while the groups
  // are assigned; it shares table 0's heap, which is idle then
  static_assert(sizeof(H[0].heap) >= 260 * sizeof(uint64_t), "packed lengths fit table 0's heap");
  uint64_t* PL = reinterpret_cast<uint64_t*>(H[0].heap);
  const int b = d.horder[blockIdx.x];
  const int tid = threadIdx.x, wv = tid >> 6, lane = lane_id();
  const BlockDesc D = d.desc[b];
  const BlockState* st = d.st + b;
  const int ng = st->n_groups, nsel = st->n_sel, s = st->s;
  const int alpha = st->n_in_use + 2;
  // the tables' current lengths, rewritten by every build (a workgroup
  // barrier orders a wave's global stores before the other waves' loads)
  uint8_t* glen = d.len + (int64_t)b * 6 * 260;
  const uint32_t* sym2 = reinterpret_cast<const uint32_t*>(d.SYM + D.off);  // 2 symbols per word
  for (int it = 0; it < 4; it++) {
    for (int v = tid; v < 260; v += kHuffThreads) {
      uint64_t w = 0;
#pragma unroll
      for (int tt = 0; tt < 6; tt++) w |= (uint64_t)glen[tt * 260 + v] << (10 * tt);
      PL[v] = w;
    }
    for (int q = tid; q < 6 * 260; q += kHuffThreads) (&rf[0][0])[q] = 0;
    if (tid == 0) s_chg = it == 0 ? 0x3f : 0;
    __syncthreads();
    // every group but a block's last holds kGSize symbols: that path runs
    // without per-symbol guards
    for (int g = tid; g < nsel; g += kHuffThreads) {
      const int gs = g * kGSize;
      const bool full = gs + kGSize <= s;
      const int ge = full ? gs + kGSize : s;
      uint32_t w[kGSize / 2];
      uint64_t acc = 0;
      if (full) {
#pragma unroll
        for (int q = 0; q < kGSize / 2; q++) w[q] = sym2[(gs >> 1) + q];
#pragma unroll
        for (int q = 0; q < kGSize; q++) acc += PL[(w[q >> 1] >> (16 * (q & 1))) & 0xFFFFu];
      } else {
#pragma unroll
        for (int q = 0; q < kGSize / 2; q++) w[q] = gs + 2 * q < ge ? sym2[(gs >> 1) + q] : 0u;
#pragma unroll
        for (int q = 0; q < kGSize; q++)
          if (gs + q < ge) acc += PL[(w[q >> 1] >> (16 * (q & 1))) & 0xFFFFu];
      }
      uint32_t bc = (uint32_t)(acc & 1023u);
      int bt = 0;
      for (int tt = 1; tt < ng; tt++) {
        const uint32_t c = (uint32_t)(acc >> (10 * tt)) & 1023u;
        if (c < bc) { bc = c; bt = tt; }
      }
      // a table whose groups are those of the previous iteration gets the
      // same counts, hence the same lengths: its build is skipped
      uint8_t* sp = d.sel + D.sel_off + g;
      if (it > 0) {
        const int old = *sp;
        if (old != bt) atomicOr(&s_chg, (1 << old) | (1 << bt));
      }
      *sp = (uint8_t)bt;
      int32_t* rt = rf[bt];
      if (full) {
#pragma unroll
        for (int q = 0; q < kGSize; q++) atomicAdd(&rt[(w[q >> 1] >> (16 * (q & 1))) & 0xFFFFu], 1);
      } else {
#pragma unroll
        for (int q = 0; q < kGSize; q++)
          if (gs + q < ge) atomicAdd(&rt[(w[q >> 1] >> (16 * (q & 1))) & 0xFFFFu], 1);
      }
    }
    __syncthreads();
    const int chg = s_chg;
    // no table changed: every later iteration repeats this one
    if (chg == 0) break;
    if (wv < ng && ((chg >> wv) & 1)) build_lengths(H[wv], rf[wv], glen + wv * 260, alpha, lane);
    __syncthreads();
  }
}

// Codes (huffman.cu:135-146, :352-362), selector MTF (compress.cu:119-137)
// and the block header bits (compress.cu:290-356) into the block's header
// words, one 256-thread workgroup per block.  Codes are canonical: thread
// (table, length) hands out consecutive codes to that length's symbols in
// symbol order.  The selector MTF runs over 256 chunks whose entering list
// state comes from a max-scan of per-chunk last occurrences.
constexpr int kSelLds = 24576;
__global__ void __launch_bounds__(256) k_huff_codes(Dev d) {
  __shared__ uint8_t SL[kSelLds];
  __shared__ uint8_t L[6][260];
  __shared__ int32_t lcount[6][32];
  __shared__ uint32_t lbase[6][32];
  __shared__ int32_t tbits[8];
  __shared__ int32_t tmp[8];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  BlockState* st = d.st + b;
  const BlockDesc D = d.desc[b];
  const int ng = st->n_groups, nsel = st->n_sel;
  const int alpha = st->n_in_use + 2;
  uint32_t pres[8];  // the block's byte-presence bits, loaded once
#pragma unroll
  for (int q = 0; q < 8; q++) pres[q] = st->present[q];
  for (int q = tid; q < 6 * 32; q += 256) (&lcount[0][0])[q] = 0;
  __syncthreads();
  for (int q = tid; q < 6 * 260; q += 256) {
    const int tb = q / 260, v = q - 260 * tb;
    const uint8_t l = d.len[(int64_t)b * 6 * 260 + q];
    L[tb][v] = l;
    if (tb < ng && v < alpha) atomicAdd(&lcount[tb][l & 31], 1);
  }
  __syncthreads();
  if (tid < ng) {
    // canonical first code of every length (lengths are 1..20)
    uint32_t vec = 0;
    for (int l = 1; l < 32; l++) {
      lbase[tid][l] = vec;
      vec = (vec + (uint32_t)lcount[tid][l]) << 1;
    }
    // coding-table bits: 5 + sum(1 + 2*|delta|)
    int cur = L[tid][0], nb = 5;
#pragma unroll 8
    for (int v = 0; v < alpha; v++) {
      const int tg = L[tid][v];
      nb += 1 + 2 * (tg > cur ? tg - cur : cur - tg);
      cur = tg;
    }
    tbits[tid] = nb;
  }
  __syncthreads();
  {
    const int tb = tid >> 5, l = tid & 31;
    if (tb < ng && l > 0 && lcount[tb][l] > 0) {
      uint32_t* code = d.code + ((int64_t)b * 6 + tb) * 260;
      uint32_t c = lbase[tb][l];
#pragma unroll 8
      for (int v = 0; v < alpha; v++)
        if (L[tb][v] == l) code[v] = c++;
    }
  }
  // selector MTF
  // the selectors, in LDS when they fit (blocks of up to ~1.2M symbols),
  // else in the block's global scratch; their MTF indexes replace them
  const uint8_t* sel = d.sel + D.sel_off;
  uint8_t* sl = nsel <= kSelLds ? SL : d.selmtf + D.sel_off;
#pragma unroll 16
  for (int i = tid; i < nsel; i += 256) sl[i] = sel[i];
  __syncthreads();
  const int per = (nsel + 255) / 256;
  const int c0 = min(nsel, tid * per), c1 = min(nsel, c0 + per);
  int lastocc[6];
#pragma unroll
  for (int q = 0; q < 6; q++) lastocc[q] = -1;
  for (int i = c0; i < c1; i++) {
    const int v = sl[i];
#pragma unroll
    for (int q = 0; q < 6; q++) lastocc[q] = v == q ? i : lastocc[q];
  }
  int ent[6];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    __shared__ int32_t xs[256];
    const int v = wg_incl_max<256, int>(lastocc[q], tmp, -1);
    xs[tid] = v;
    __syncthreads();
    ent[q] = tid > 0 ? xs[tid - 1] : -1;
    __syncthreads();
    // never-seen ids rank after seen ones, in ascending order (initial list)
    if (ent[q] < 0) ent[q] = -1000000000 - q;
  }
  // the list as 6 nibbles (nibble q = the table id at list position q)
  // (tables in use ordered by descending entering position, which are
  // distinct; the others keep their places)
  uint32_t lst = 0;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    int pos = q;
    if (q < ng) {
      pos = 0;
#pragma unroll
      for (int r = 0; r < 6; r++) pos += (r < ng && ent[r] > ent[q]) ? 1 : 0;
    }
    lst |= (uint32_t)q << (4 * pos);
  }
  int my_bits = 0;
  for (int i = c0; i < c1; i++) {
    const uint32_t v = sl[i];
    // position of v: the lowest zero nibble of lst ^ v-in-every-nibble
    const uint32_t x = lst ^ (v * 0x111111u);
    const int j = __builtin_ctz((x - 0x111111u) & ~x & 0x888888u) >> 2;
    const uint32_t low = (1u << (4 * j)) - 1u;
    lst = (lst & ~((low << 4) | 15u)) | ((lst & low) << 4) | v;
    sl[i] = (uint8_t)j;
    my_bits += j + 1;
  }
  // header layout
  int nused = 0;
  uint32_t l1 = 0;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    uint32_t wv = pres[r >> 1];
    uint32_t half = (r & 1) ? (wv >> 16) : (wv & 0xFFFF);
    if (half) { l1 |= 1u << (15 - r); nused++; }
  }
  const int64_t fixed = 48 + 32 + 1 + 24 + 16 + 16 * nused + 3 + 15;
  int sel_total;
  const int sb_ex = wg_excl_sum<256, int>(my_bits, tmp, &sel_total);
  int tb_total = 0;
  for (int q = 0; q < ng; q++) tb_total += tbits[q];
  uint32_t* hw = d.hdr + (int64_t)b * d.hdr_stride;
  const int64_t hdr_bits = fixed + sel_total + tb_total;
  // fixed part
  if (tid == 0) {
    BitOut o;
    o.init(hw, 0, fixed, false);
    o.put(0x314159u, 24); o.put(0x265359u, 24);
    o.put(st->crc, 32);
    o.put(0, 1);
    o.put((uint32_t)st->orig_ptr, 24);
    o.put(l1, 16);
#pragma unroll
    for (int r = 0; r < 16; r++) {
      if (!((l1 >> (15 - r)) & 1u)) continue;
      uint32_t wv = pres[r >> 1];
      uint32_t half = (r & 1) ? (wv >> 16) : (wv & 0xFFFF);
      uint32_t l2 = 0;
      for (int q = 0; q < 16; q++) if ((half >> q) & 1u) l2 |= 1u << (15 - q);
      o.put(l2, 16);
    }
    o.put((uint32_t)ng, 3);
    o.put((uint32_t)nsel, 15);
    o.finish();
    st->hdr_bits = hdr_bits;
  }
  // selectors: this thread's chunk owns [fixed + sb_ex, + my_bits)
  {
    BitOut o;
    const int64_t a = fixed + sb_ex;
    o.init(hw, a, a + my_bits, false);
    for (int i = c0; i < c1; i++) {
      const int j = sl[i];
      o.put(((1u << j) - 1) << 1, j + 1);  // j ones then a zero
    }
    o.finish();
  }
  // coding tables: thread q < ng owns its table's range
  if (tid < ng) {
    int tb_before = 0;
    for (int q = 0; q < tid; q++) tb_before += tbits[q];
    BitOut o;
    const int64_t a = fixed + sel_total + tb_before;
    o.init(hw, a, a + tbits[tid], false);
    int cur = L[tid][0];
    o.put((uint32_t)cur, 5);
#pragma unroll 4
    for (int v = 0; v < alpha; v++) {
      const int tg = L[tid][v];
      const int dd = tg - cur, ad = dd < 0 ? -dd : dd;
      if (ad <= 15) {
        // |dd| steps "10" (up) or "11" (down), then "0", in one put
        const uint32_t m2 = (1u << (2 * ad)) - 1u;
        o.put((dd >= 0 ? 0xAAAAAAAAu & m2 : m2) << 1, 2 * ad + 1);
      } else {
        while (cur < tg) { o.put(2, 2); cur++; }
        while (cur > tg) { o.put(3, 2); cur--; }
        o.put(0, 1);
      }
      cur = tg;
    }
    o.finish();
  }
}

// ------------------------------------------------------------ payload
// bits per 50-symbol group
// Payload bits of every capacity tile's selector groups (tile k holds the
// groups that start in its 4096 symbols): the symbols staged in LDS, code
// lengths summed per thread, then per workgroup.
__global__ void __launch_bounds__(256) k_group_bits(Dev d) {
  __shared__ uint8_t L[6][260];
  __shared__ uint8_t T[kTile / kGSize + 2];
  __shared__ int32_t tmp[8];
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int s = d.st[b].s;
  const int k = t - D.tile_c0;
  const int g0 = (k * kTile + kGSize - 1) / kGSize;
  const int g1 = min(((k + 1) * kTile + kGSize - 1) / kGSize, (s + kGSize - 1) / kGSize);
  const int tid = threadIdx.x;
  if (g0 >= g1) {
    if (tid == 0) d.gbits[D.sel_off + k] = 0;
    return;
  }
  constexpr int kSyms = (kTile / kGSize + 2) * kGSize;
  constexpr int kTab = (6 * 260 + 255) / 256, kSw = (kSyms / 2 + 255) / 256;
  static_assert(kTile / kGSize + 2 <= 256, "one selector per thread");
  static_assert(kGSize % 2 == 0, "a symbol pair lies in one group");
  const int sbase = g0 * kGSize, n = min(g1 * kGSize, s) - sbase;
  // all loads in flight at once: lengths and selectors to LDS, symbol pairs
  // (sbase is even, the block's symbols start 4-byte aligned) to registers
  const uint8_t* len = d.len + (int64_t)b * 6 * 260;
  const uint32_t* sym2 = reinterpret_cast<const uint32_t*>(d.SYM + D.off + sbase);
  const int nw2 = (n + 1) >> 1;
  uint8_t lv[kTab];
  uint32_t sv[kSw];
#pragma unroll
  for (int j = 0; j < kTab; j++) {
    const int q = tid + 256 * j;
    lv[j] = q < 6 * 260 ? len[q] : (uint8_t)0;
  }
#pragma unroll
  for (int j = 0; j < kSw; j++) {
    const int q = tid + 256 * j;
    sv[j] = q < nw2 ? sym2[q] : 0u;
  }
  const uint8_t tv = tid < g1 - g0 ? d.sel[D.sel_off + g0 + tid] : (uint8_t)0;
#pragma unroll
  for (int j = 0; j < kTab; j++) {
    const int q = tid + 256 * j;
    if (q < 6 * 260) (&L[0][0])[q] = lv[j];
  }
  if (tid < g1 - g0) T[tid] = tv;
  __syncthreads();
  int sum = 0;
#pragma unroll
  for (int j = 0; j < kSw; j++) {
    const int q = tid + 256 * j;
    if (q < nw2) {
      const int tt = T[(2 * q) / kGSize];
      sum += L[tt][sv[j] & 0xFFFFu];
      if (2 * q + 1 < n) sum += L[tt][sv[j] >> 16];
    }
  }
  int tot;
  (void)wg_excl_sum<256, int>(sum, tmp, &tot);
  if (tid == 0) d.gbits[D.sel_off + k] = tot;
}

// per block: exclusive scan of the tiles' payload bits -> each tile's first
// bit, the block's bit count, zeroed payload words
__global__ void __launch_bounds__(256) k_group_scan(Dev d) {
  __shared__ int32_t tmp[8];
  const int b = blockIdx.x;
  const BlockDesc D = d.desc[b];
  const int nt = tiles_for(D.cap);
  int run = 0;
  for (int base = 0; base < nt; base += 256) {
    const int k = base + threadIdx.x;
    const int v = k < nt ? d.gbits[D.sel_off + k] : 0;
    int tot;
    const int ex = wg_excl_sum<256, int>(v, tmp, &tot);
    if (k < nt) d.gbits[D.sel_off + k] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) d.st[b].bits = run;
  const int64_t nw = ((int64_t)run + 31) / 32 + 1;
  for (int64_t q = threadIdx.x; q < nw; q += 256) d.pay[D.w_off + q] = 0;
}

// Payload bits of one capacity tile's selector groups (huffman_encode.cu:
// 15-71): each thread takes up to kPackRun consecutive symbols, a workgroup
// scan of their code lengths gives every symbol's bit offset, the codes are
// ORed into the tile's words staged in LDS, and the words go out coalesced
// (plain stores inside the tile's bit range, atomicOr on the two edge words
// it shares with the neighbouring tiles).
constexpr int kPackGroups = kTile / kGSize + 2;  // selector groups a tile can start
constexpr int kPackSyms = kPackGroups * kGSize;
constexpr int kPackRun = (kPackSyms + 255) / 256;
constexpr int kPackWords = kPackSyms * 17 / 32 + 2;

__global__ void __launch_bounds__(256) k_pack(Dev d) {
  __shared__ uint32_t CL[6][260];  // code << 5 | length
  __shared__ __align__(4) uint16_t S[kPackSyms + 2];
  __shared__ uint8_t T[kPackGroups];
  __shared__ uint32_t W[kPackWords];
  __shared__ int32_t tmp[8];
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int s = d.st[b].s;
  const int k = t - D.tile_c0;
  const int g0 = (k * kTile + kGSize - 1) / kGSize;
  const int nsel = (s + kGSize - 1) / kGSize;
  const int g1 = min(((k + 1) * kTile + kGSize - 1) / kGSize, nsel);
  if (g0 >= g1) return;
  const int tid = threadIdx.x;
  const int sbase = g0 * kGSize, n = min(g1 * kGSize, s) - sbase;
  // all loads of the tables, symbols (two per word: sbase is even and the
  // block's symbols start 4-byte aligned) and selectors in flight at once
  {
    constexpr int kTab = (6 * 260 + 255) / 256, kSw = (kPackSyms / 2 + 255) / 256;
    const uint32_t* code = d.code + (int64_t)b * 6 * 260;
    const uint8_t* len = d.len + (int64_t)b * 6 * 260;
    const uint32_t* sym2 = reinterpret_cast<const uint32_t*>(d.SYM + D.off + sbase);
    const int nw2 = (n + 1) >> 1;
    uint32_t cv[kTab], sv[kSw];
    uint8_t lv[kTab];
#pragma unroll
    for (int j = 0; j < kTab; j++) {
      const int q = tid + 256 * j;
      cv[j] = q < 6 * 260 ? code[q] : 0u;
      lv[j] = q < 6 * 260 ? len[q] : (uint8_t)0;
    }
#pragma unroll
    for (int j = 0; j < kSw; j++) {
      const int q = tid + 256 * j;
      sv[j] = q < nw2 ? sym2[q] : 0u;
    }
    const uint8_t tv = tid < g1 - g0 ? d.sel[D.sel_off + g0 + tid] : (uint8_t)0;
#pragma unroll
    for (int j = 0; j < kTab; j++) {
      const int q = tid + 256 * j;
      if (q < 6 * 260) (&CL[0][0])[q] = (cv[j] << 5) | lv[j];
    }
#pragma unroll
    for (int j = 0; j < kSw; j++) {
      const int q = tid + 256 * j;
      if (q < nw2) reinterpret_cast<uint32_t*>(S)[q] = sv[j];
    }
    if (tid < g1 - g0) T[tid] = tv;
  }
  static_assert(kPackGroups <= 256, "one selector per thread");
  const int64_t a = d.gbits[D.sel_off + k];  // the tile's bit range [a, e)
  const int64_t e = g1 < nsel ? d.gbits[D.sel_off + k + 1] : d.st[b].bits;
  const int64_t w0 = a >> 5;
  const int nw = (int)(((e + 31) >> 5) - w0);
  for (int q = tid; q < nw; q += 256) W[q] = 0u;
  __syncthreads();
  const int i0 = tid * kPackRun, cnt = max(0, min(kPackRun, n - i0));
  uint32_t cl[kPackRun];
  int sum = 0;
#pragma unroll
  for (int j = 0; j < kPackRun; j++) {
    cl[j] = 0;
    if (j < cnt) {
      const int i = i0 + j;
      cl[j] = CL[T[i / kGSize]][S[i]];
      sum += (int)(cl[j] & 31u);
    }
  }
  const int off = wg_excl_sum<256, int>(sum, tmp, nullptr) + (int)(a & 31);
  // The thread's bits [off, off + sum) gather in a 64-bit register; each
  // completed word is stored, by atomicOr only where a neighbour's bits share
  // it (the first word when off is not word-aligned, and the last, partial
  // word).
  {
    const int wfirst = off >> 5;
    int wi = wfirst, nb = off & 31;
    uint64_t buf = 0;
#pragma unroll
    for (int j = 0; j < kPackRun; j++) {
      const int l = (int)(cl[j] & 31u);
      buf = (buf << l) | (uint64_t)(cl[j] >> 5);
      nb += l;
      if (nb >= 32) {
        const uint32_t word = (uint32_t)(buf >> (nb - 32));
        if (wi == wfirst && (off & 31)) atomicOr(&W[wi], word);
        else W[wi] = word;
        wi++;
        nb -= 32;
        buf &= (1ull << nb) - 1ull;
      }
    }
    if (nb > 0 && sum > 0) atomicOr(&W[wi], (uint32_t)(buf << (32 - nb)));
  }
  __syncthreads();
  uint32_t* pay = d.pay + D.w_off;
  for (int q = tid; q < nw; q += 256) {
    const int64_t wi = w0 + q;
    const uint32_t v = W[q];
    if (wi * 32 >= a && wi * 32 + 32 <= e) pay[wi] = v;
    else if (v) atomicOr(&pay[wi], v);
  }
}

// ------------------------------------------------------------ assembly
// One workgroup per stream: bit offset of every block inside its stream (an
// exclusive scan of header + payload bits after the 32-bit stream header),
// the stream's bit length and its combined CRC.  The combined CRC folds
// c = rotl(c, 1) ^ crc_q over the blocks, i.e. the XOR of every block CRC
// rotated by the number of blocks after it.
__global__ void __launch_bounds__(256) k_layout_blocks(Dev d) {
  __shared__ int64_t tmp64[4];
  __shared__ uint32_t xr[4];
  const int si = blockIdx.x;
  const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
  StreamState S = d.ss[si];
  const bool seg = S.lead >= 0;
  int64_t run = seg ? S.lead : 32;
  // blocks before a segment: c = rotl(c, 1) ^ crc_q continues from crc_in,
  // which ends up rotated by this segment's block count
  const int rin = S.n_blk & 31;
  uint32_t crc = seg ? (rin ? (S.crc_in << rin) | (S.crc_in >> (32 - rin)) : S.crc_in) : 0u;
  for (int q0 = 0; q0 < S.n_blk; q0 += 256) {
    const int q = q0 + tid;
    int64_t nb = 0;
    uint32_t c = 0;
    BlockState* st = d.st + S.first_blk + q;
    if (q < S.n_blk) {
      nb = st->hdr_bits + st->bits;
      const int r = (S.n_blk - 1 - q) & 31;
      c = r ? (st->crc << r) | (st->crc >> (32 - r)) : st->crc;
    }
    const int64_t inc = wave_incl_sum(nb);
    if (lane == 63) tmp64[wv] = inc;
    for (int o = 32; o >= 1; o >>= 1) c ^= (uint32_t)__shfl_xor((int)c, o, 64);
    if (lane == 0) xr[wv] = c;
    __syncthreads();
    int64_t pre = 0;
    for (int w = 0; w < wv; w++) pre += tmp64[w];
    if (q < S.n_blk) st->bit_off = run + pre + inc - nb;
    run += tmp64[0] + tmp64[1] + tmp64[2] + tmp64[3];
    crc ^= xr[0] ^ xr[1] ^ xr[2] ^ xr[3];
    __syncthreads();
  }
  if (tid == 0) {
    S.bits = seg ? run : run + 48 + 32;
    S.crc = crc;
    S.bytes = (S.bits + 7) / 8;
    d.ss[si] = S;
  }
}

// One workgroup: the streams' 4-byte aligned places in the output (a scan
// of their word counts) and their header words.
__global__ void __launch_bounds__(256) k_layout(Dev d) {
  __shared__ int64_t tmp64[8];
  uint32_t* out = reinterpret_cast<uint32_t*>(d.out);
  int64_t run = 0;
  for (int base = 0; base < d.nstream; base += 256) {
    const int si = base + threadIdx.x;
    int64_t words = 0;
    StreamState S{};
    if (si < d.nstream) {
      S = d.ss[si];
      words = (S.bits + 31) / 32;
    }
    int64_t inc = wave_incl_sum(words);
    if (lane_id() == 63) tmp64[threadIdx.x >> 6] = inc;
    __syncthreads();
    int64_t pre = 0;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) pre += tmp64[w];
    int64_t tot = tmp64[0] + tmp64[1] + tmp64[2] + tmp64[3];
    __syncthreads();
    if (si < d.nstream) {
      S.byte_off = (run + pre + inc - words) * 4;
      d.ss[si] = S;
      uint32_t* w0 = out + S.byte_off / 4;
      if (S.lead < 0) {
        w0[0] = __builtin_bswap32(0x425A6830u + (uint32_t)S.level);
        // the words shared at the end-of-stream marker
        const int64_t ta = S.bits - 80;
        w0[ta >> 5] = 0;
        w0[(S.bits - 1) >> 5] = 0;
      } else if (S.bits > 0) {
        w0[0] = 0;  // the lead bits stay zero (an empty segment has no words)
      }
    }
    run += tot;
  }
}

// Zeroes the words a block shares with its neighbours (first and last word,
// and the one at its header/payload seam), which the placement ORs into.
__global__ void __launch_bounds__(256) k_seams(Dev d) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= d.nblk) return;
  const BlockState* st = d.st + b;
  const StreamState S = d.ss[d.desc[b].stream];
  uint32_t* w0 = reinterpret_cast<uint32_t*>(d.out) + S.byte_off / 4;
  const int64_t a = st->bit_off, e = a + st->hdr_bits + st->bits;
  w0[a >> 5] = 0;
  w0[(e - 1) >> 5] = 0;
  w0[(a + st->hdr_bits) >> 5] = 0;
}

// Copies `nbits` bits from aligned source words into the stream at bit
// `dst` (stream-relative), shared edge words by atomicOr.
__device__ inline void place_bits(uint32_t* w0, const uint32_t* src, int64_t dst, int64_t nbits,
                                  int tid, int nth) {
  if (nbits <= 0) return;
  const int sh = (int)(dst & 31);
  const int64_t wa = dst >> 5, we = (dst + nbits - 1) >> 5;
  const int64_t nsrcw = (nbits + 31) / 32;
  for (int64_t wi = wa + tid; wi <= we; wi += nth) {
    // destination word wi gets source bits [(wi-wa)*32 - sh, +32)
    int64_t j = wi - wa;
    uint32_t hi = (j - 1 >= 0 && j - 1 < nsrcw) ? src[j - 1] : 0u;
    uint32_t lo = j < nsrcw ? src[j] : 0u;
    uint32_t v = sh == 0 ? lo : ((hi << (32 - sh)) | (lo >> sh));
    // mask off bits outside [dst, dst+nbits)
    int64_t wbit0 = wi * 32;
    uint32_t mask = 0xFFFFFFFFu;
    if (wbit0 < dst) mask &= 0xFFFFFFFFu >> (dst - wbit0);
    int64_t endb = dst + nbits;
    if (wbit0 + 32 > endb) mask &= ~(0xFFFFFFFFu >> (endb - wbit0));
    v &= mask;
    if (mask == 0xFFFFFFFFu) w0[wi] = __builtin_bswap32(v);
    else if (v) atomicOr(&w0[wi], __builtin_bswap32(v));
  }
}

// One workgroup per capacity tile; the tiles of a block share its payload
// words round-robin (256 consecutive words per workgroup and step).
__global__ void __launch_bounds__(256) k_place(Dev d) {
  const int t = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int b = d.tc_blk[t];
  const BlockDesc D = d.desc[b];
  const int k = t - D.tile_c0;
  const int ntk = tiles_for(D.cap);
  const BlockState* st = d.st + b;
  const StreamState S = d.ss[D.stream];
  uint32_t* w0 = reinterpret_cast<uint32_t*>(d.out) + S.byte_off / 4;
  if (k == 0) place_bits(w0, d.hdr + (int64_t)b * d.hdr_stride, st->bit_off, st->hdr_bits, threadIdx.x, 256);
  place_bits(w0, d.pay + D.w_off, st->bit_off + st->hdr_bits, st->bits, k * 256 + threadIdx.x, ntk * 256);
}

__global__ void __launch_bounds__(64) k_trailer(Dev d) {
  const int si = blockIdx.x * 64 + threadIdx.x;
  if (si >= d.nstream) return;
  const StreamState S = d.ss[si];
  if (S.lead >= 0) return;  // segments end without the trailer
  uint32_t* w0 = reinterpret_cast<uint32_t*>(d.out) + S.byte_off / 4;
  uint32_t src[3] = {0x17724538u, (0x5090u << 16) | (S.crc >> 16), (S.crc & 0xFFFFu) << 16};
  place_bits(w0, src, S.bits - 80, 80, 0, 1);
}

void launch_huffman(const Dev& d, hipStream_t s) {
  k_huff_init<<<d.nblk, 64, 0, s>>>(d);
  k_huff_order<<<1, 256, 0, s>>>(d);
  probe_begin("k_huff_build", s);
  k_huff_build<<<d.nblk, kHuffThreads, 0, s>>>(d);
  probe_end("k_huff_build", s);
  k_huff_codes<<<d.nblk, 256, 0, s>>>(d);
}

void launch_pack(const Dev& d, hipStream_t s) {
  k_group_bits<<<d.ntile_c, 256, 0, s>>>(d);
  k_group_scan<<<d.nblk, 256, 0, s>>>(d);
  probe_begin("k_pack", s);
  k_pack<<<d.ntile_c, 256, 0, s>>>(d);
  probe_end("k_pack", s);
}

void launch_assemble(const Dev& d, hipStream_t s) {
  if (d.nstream == 0) return;
  k_layout_blocks<<<d.nstream, 256, 0, s>>>(d);
  k_layout<<<1, 256, 0, s>>>(d);
  if (d.nblk > 0) {
    k_seams<<<(d.nblk + 255) / 256, 256, 0, s>>>(d);
    k_place<<<d.ntile_c, 256, 0, s>>>(d);
  }
  k_trailer<<<(d.nstream + 63) / 64, 64, 0, s>>>(d);
}

}  // namespace bzg
