/*
 * oracle.c - CPU restatement of the reference bzip2 block-compression path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See oracle.h
 * for what pins each function.  Citations are to the reference repository
 * SebastianBoteroLeonik/bzip2gpu.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ CRC */
/* src/compression/crc32.cpp:6-29: MSB-first table CRC, polynomial
 * 0x04C11DB7, register initialised to all ones, result inverted. */
static uint32_t crc_table[256];
static pthread_once_t crc_once = PTHREAD_ONCE_INIT;

static void crc_init(void) {
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t c = i << 24;
    for (int j = 0; j < 8; j++) c = (c & 0x80000000u) ? (c << 1) ^ 0x04C11DB7u : (c << 1);
    crc_table[i] = c;
  }
}

uint32_t orc_crc32_update(uint32_t reg, const uint8_t *data, size_t len) {
  pthread_once(&crc_once, crc_init);
  for (size_t i = 0; i < len; i++) reg = (reg << 8) ^ crc_table[(reg >> 24) ^ data[i]];
  return reg;
}

uint32_t orc_crc32(const uint8_t *data, size_t len) {
  return ~orc_crc32_update(0xFFFFFFFFu, data, len);
}

/* ----------------------------------------------------------------- RLE1 */
/* src/compression/rle1.cu:13-58 (RunSizeCalc + write_bzip2_rle1_kernel):
 * each maximal run of c equal bytes becomes floor(c/259) groups of
 * "x x x x 255", then a remainder r = c % 259 written as r literals when
 * r < 4, else "x x x x (r-4)". */
int32_t orc_rle1(const uint8_t *in, int32_t n, uint8_t *out) {
  int32_t o = 0, i = 0;
  while (i < n) {
    int32_t j = i + 1;
    while (j < n && in[j] == in[i]) j++;
    int32_t c = j - i;
    uint8_t x = in[i];
    int32_t full = c / 259, rem = c % 259;
    for (int32_t q = 0; q < full; q++) {
      out[o++] = x; out[o++] = x; out[o++] = x; out[o++] = x; out[o++] = 255;
    }
    if (rem < 4) {
      for (int32_t q = 0; q < rem; q++) out[o++] = x;
    } else {
      out[o++] = x; out[o++] = x; out[o++] = x; out[o++] = x;
      out[o++] = (uint8_t)(rem - 4);
    }
    i = j;
  }
  return o;
}

/* ------------------------------------------------------------------ BWT */
static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* Sorted cyclic rotations (src/compression/fbwt.cu:13-95: prefix doubling
 * over (rank[SA[i]], rank[(SA[i]+k) mod m]) pairs, stable from the identity,
 * so identical rotations end up in ascending index order,
 * tests/compression/fbwt.cu:22).  Restated as class doubling with stable
 * counting sorts; identical rotations are put in index order at the end. */
int32_t orc_bwt(const uint8_t *s, int32_t n, int32_t *p) {
  if (n <= 0) return 0;
  int32_t *c = (int32_t *)malloc(sizeof(int32_t) * n);
  int32_t *cn = (int32_t *)malloc(sizeof(int32_t) * n);
  int32_t *pn = (int32_t *)malloc(sizeof(int32_t) * n);
  int32_t cnt_len = n > 256 ? n : 256;
  int32_t *cnt = (int32_t *)calloc(cnt_len, sizeof(int32_t));
  for (int32_t i = 0; i < n; i++) cnt[s[i]]++;
  for (int32_t i = 1; i < 256; i++) cnt[i] += cnt[i - 1];
  for (int32_t i = n - 1; i >= 0; i--) p[--cnt[s[i]]] = i;
  int32_t classes = 1;
  c[p[0]] = 0;
  for (int32_t i = 1; i < n; i++) {
    if (s[p[i]] != s[p[i - 1]]) classes++;
    c[p[i]] = classes - 1;
  }
  int32_t rounds = 0;
  for (int64_t h = 1; h < n && classes < n; h <<= 1) {
    rounds++;
    for (int32_t i = 0; i < n; i++) {
      int64_t v = (int64_t)p[i] - h;
      v %= n;
      if (v < 0) v += n;
      pn[i] = (int32_t)v;
    }
    memset(cnt, 0, sizeof(int32_t) * classes);
    for (int32_t i = 0; i < n; i++) cnt[c[pn[i]]]++;
    for (int32_t i = 1; i < classes; i++) cnt[i] += cnt[i - 1];
    for (int32_t i = n - 1; i >= 0; i--) p[--cnt[c[pn[i]]]] = pn[i];
    int32_t nc = 1;
    cn[p[0]] = 0;
    for (int32_t i = 1; i < n; i++) {
      int32_t a0 = c[p[i]], a1 = c[(int32_t)(((int64_t)p[i] + h) % n)];
      int32_t b0 = c[p[i - 1]], b1 = c[(int32_t)(((int64_t)p[i - 1] + h) % n)];
      if (a0 != b0 || a1 != b1) nc++;
      cn[p[i]] = nc - 1;
    }
    int32_t *t = c; c = cn; cn = t;
    classes = nc;
  }
  /* identical rotations (periodic blocks): ascending start index */
  if (classes < n) {
    int32_t i = 0;
    while (i < n) {
      int32_t j = i + 1;
      while (j < n && c[p[j]] == c[p[i]]) j++;
      if (j - i > 1) qsort(p + i, (size_t)(j - i), sizeof(int32_t), cmp_i32);
      i = j;
    }
  }
  free(c); free(cn); free(pn); free(cnt);
  return rounds;
}

/* ------------------------------------------------------------------ MTF */
static void mtf_core(const uint8_t *L, int32_t m, const uint8_t *present,
                     uint8_t *out) {
  uint8_t list[256];
  int32_t k = 0;
  for (int32_t i = 0; i < 256; i++)
    if (present[i]) list[k++] = (uint8_t)i;
  for (int32_t i = 0; i < m; i++) {
    uint8_t c = L[i];
    int32_t pos = 0;
    while (list[pos] != c) pos++;
    out[i] = (uint8_t)pos;
    memmove(list + 1, list, (size_t)pos);
    list[0] = c;
  }
}

/* src/compression/fmtf.cu:40-97 and :147-230: L[i] = in[(SA[i]-1) mod m];
 * initial list = present bytes of the RLE1 block in ascending order
 * (fmtf.cu:180-210); output = list position; orig_ptr = first i with
 * SA[i] == 0 (fmtf.cu:222-225). */
int32_t orc_mtf(const uint8_t *in, const int32_t *sa, int32_t m, uint8_t *out,
                int32_t *orig_ptr, uint8_t *present) {
  uint8_t pres[256];
  memset(pres, 0, sizeof pres);
  for (int32_t i = 0; i < m; i++) pres[in[i]] = 1;
  int32_t n_in_use = 0;
  for (int32_t i = 0; i < 256; i++) n_in_use += pres[i];
  if (present) memcpy(present, pres, 256);
  uint8_t *L = (uint8_t *)malloc((size_t)m + 1);
  int32_t op = 0;
  int found = 0;
  for (int32_t i = 0; i < m; i++) {
    int32_t v = sa[i];
    L[i] = v == 0 ? in[m - 1] : in[v - 1];
    if (v == 0 && !found) { op = i; found = 1; }
  }
  mtf_core(L, m, pres, out);
  free(L);
  if (orig_ptr) *orig_ptr = op;
  return n_in_use;
}

/* tests/compression/fmtf.cu:48-74 */
void orc_mtf_plain(const uint8_t *in, int32_t m, uint8_t *out) {
  uint8_t pres[256];
  memset(pres, 0, sizeof pres);
  for (int32_t i = 0; i < m; i++) pres[in[i]] = 1;
  mtf_core(in, m, pres, out);
}

/* ----------------------------------------------------------------- RLE2 */
/* src/compression/rle2.cu:11-77 (per-chunk form of the sequential rule,
 * spec tests/compression/rle2.cu:9-33): nonzero v -> v+1; a zero run of
 * length k -> digits "k--, emit k&1 (RUNB=1 / RUNA=0), k >>= 1" until k==0;
 * EOB = n_in_use + 1 appended. */
int32_t orc_rle2(const uint8_t *in, int32_t m, int32_t n_in_use, uint16_t *out) {
  int32_t o = 0, i = 0;
  while (i < m) {
    if (in[i] != 0) {
      out[o++] = (uint16_t)(in[i] + 1);
      i++;
    } else {
      int32_t k = 0;
      while (i < m && in[i] == 0) { k++; i++; }
      while (k > 0) {
        k--;
        out[o++] = (uint16_t)(k & 1);
        k >>= 1;
      }
    }
  }
  out[o++] = (uint16_t)(n_in_use + 1);
  return o;
}

/* -------------------------------------------------------------- Huffman */
#define WEIGHTOF(z) ((z) & 0xffffff00)
#define DEPTHOF(z) ((z) & 0x000000ff)
#define MYMAX(a, b) ((a) > (b) ? (a) : (b))

/* src/compression/huffman.cu:51-124 (the libbz2 hbMakeCodeLengths method):
 * binary heap keyed by weight = freq<<8 | depth, zero frequency counted as
 * 1, and on any length > max_len every weight w becomes 1 + (w>>8)/2 and
 * the build is repeated. */
void orc_make_code_lengths(const int32_t *freq, uint8_t *len, int32_t alpha,
                           int32_t max_len) {
  int32_t heap[ORC_MAX_ALPHA + 2];
  int32_t weight[ORC_MAX_ALPHA * 2];
  int32_t parent[ORC_MAX_ALPHA * 2];
  for (int32_t i = 0; i < alpha; i++) weight[i + 1] = (freq[i] == 0 ? 1 : freq[i]) << 8;
  for (;;) {
    int32_t n_nodes = alpha, n_heap = 0;
    heap[0] = 0; weight[0] = 0; parent[0] = -2;
    for (int32_t i = 1; i <= alpha; i++) {
      parent[i] = -1;
      n_heap++;
      heap[n_heap] = i;
      /* sift up */
      int32_t zz = n_heap, tmp = heap[zz];
      while (weight[tmp] < weight[heap[zz >> 1]]) { heap[zz] = heap[zz >> 1]; zz >>= 1; }
      heap[zz] = tmp;
    }
    while (n_heap > 1) {
      int32_t pick[2];
      for (int q = 0; q < 2; q++) {
        pick[q] = heap[1];
        heap[1] = heap[n_heap];
        n_heap--;
        /* sift down from the root */
        int32_t zz = 1, tmp = heap[zz];
        for (;;) {
          int32_t yy = zz << 1;
          if (yy > n_heap) break;
          if (yy < n_heap && weight[heap[yy + 1]] < weight[heap[yy]]) yy++;
          if (weight[tmp] < weight[heap[yy]]) break;
          heap[zz] = heap[yy];
          zz = yy;
        }
        heap[zz] = tmp;
      }
      n_nodes++;
      parent[pick[0]] = parent[pick[1]] = n_nodes;
      weight[n_nodes] = (WEIGHTOF(weight[pick[0]]) + WEIGHTOF(weight[pick[1]])) |
                        (1 + MYMAX(DEPTHOF(weight[pick[0]]), DEPTHOF(weight[pick[1]])));
      parent[n_nodes] = -1;
      n_heap++;
      heap[n_heap] = n_nodes;
      int32_t zz = n_heap, tmp = heap[zz];
      while (weight[tmp] < weight[heap[zz >> 1]]) { heap[zz] = heap[zz >> 1]; zz >>= 1; }
      heap[zz] = tmp;
    }
    int too_long = 0;
    for (int32_t i = 1; i <= alpha; i++) {
      int32_t j = 0, k = i;
      while (parent[k] >= 0) { k = parent[k]; j++; }
      len[i - 1] = (uint8_t)j;
      if (j > max_len) too_long = 1;
    }
    if (!too_long) break;
    for (int32_t i = 1; i <= alpha; i++) {
      int32_t j = weight[i] >> 8;
      j = 1 + (j / 2);
      weight[i] = j << 8;
    }
  }
}

/* src/compression/huffman.cu:148-185: initial split of the alphabet into
 * n_groups contiguous ranges of roughly equal frequency; range g gets cost 0
 * for its symbols and 15 for all others.  The scan bound is the reference's
 * `ge < alphabet_size` (:160); it only differs from libbz2's alphaSize-1 when
 * the frequencies do not sum to s, which zeroed counting rules out. */
static void initial_assignment(int32_t n_groups, int32_t s, int32_t alpha,
                               const int32_t *freq /* alpha+1 entries */,
                               uint8_t len[ORC_MAX_GROUPS][ORC_MAX_ALPHA]) {
  int32_t gs = 0, rem = s;
  for (int32_t part = n_groups; part > 0; part--) {
    int32_t t = rem / part, ge = gs - 1, a = 0;
    while (a < t && ge < alpha) { ge++; a += freq[ge]; }
    if (ge > gs && part != n_groups && part != 1 && ((n_groups - part) % 2 == 1)) {
      a -= freq[ge];
      ge--;
    }
    for (int32_t v = 0; v < alpha; v++) len[part - 1][v] = (v >= gs && v <= ge) ? 0 : 15;
    gs = ge + 1;
    rem -= a;
  }
}

/* src/compression/huffman.cu:135-146 */
static void assign_codes(int32_t *code, const uint8_t *len, int32_t min_len,
                         int32_t max_len, int32_t alpha) {
  int32_t vec = 0;
  for (int32_t l = min_len; l <= max_len; l++) {
    for (int32_t v = 0; v < alpha; v++)
      if (len[v] == l) code[v] = vec++;
    vec <<= 1;
  }
}

/* src/compression/huffman.cu:256-364 with the frequency accumulator zeroed
 * (the reference allocates it uninitialised at :281-286; its own test zeroes
 * it, tests/compression/huffman.cu:26, and libbz2 counts from zero). */
int32_t orc_huff_build(const uint16_t *sym, int32_t s, int32_t alphabet_size,
                       uint8_t len[ORC_MAX_GROUPS][ORC_MAX_ALPHA],
                       int32_t code[ORC_MAX_GROUPS][ORC_MAX_ALPHA],
                       uint8_t *sel, int32_t *n_groups_out) {
  int32_t alpha = alphabet_size + 2;
  for (int g = 0; g < ORC_MAX_GROUPS; g++)
    for (int v = 0; v < alpha; v++) len[g][v] = 15;
  int32_t n_groups = s < 200 ? 2 : s < 600 ? 3 : s < 1200 ? 4 : s < 2400 ? 5 : 6;
  int32_t freq[ORC_MAX_ALPHA + 1];
  memset(freq, 0, sizeof freq);
  for (int32_t i = 0; i < s; i++) freq[sym[i]]++;
  initial_assignment(n_groups, s, alpha, freq, len);
  int32_t n_sel = (s + ORC_G_SIZE - 1) / ORC_G_SIZE;
  int32_t rfreq[ORC_MAX_GROUPS][ORC_MAX_ALPHA];
  for (int iter = 0; iter < 4; iter++) {
    memset(rfreq, 0, sizeof rfreq);
    for (int32_t g = 0; g < n_sel; g++) {
      int32_t gs = g * ORC_G_SIZE, ge = gs + ORC_G_SIZE - 1;
      if (ge >= s) ge = s - 1;
      uint32_t cost[ORC_MAX_GROUPS] = {0, 0, 0, 0, 0, 0};
      for (int32_t i = gs; i <= ge; i++)
        for (int t = 0; t < n_groups; t++) cost[t] += len[t][sym[i]];
      /* first strict minimum (huffman.cu:220-231) */
      uint32_t best_cost = 0xFFFFFFFFu;
      int best = -1;
      for (int t = 0; t < n_groups; t++)
        if (cost[t] < best_cost) { best_cost = cost[t]; best = t; }
      sel[g] = (uint8_t)best;
      for (int32_t i = gs; i <= ge; i++) rfreq[best][sym[i]]++;
    }
    for (int t = 0; t < n_groups; t++) orc_make_code_lengths(rfreq[t], len[t], alpha, 17);
  }
  for (int t = 0; t < n_groups; t++) {
    int32_t mn = 32, mx = 0;
    for (int32_t v = 0; v < alpha; v++) {
      if (len[t][v] > mx) mx = len[t][v];
      if (len[t][v] < mn) mn = len[t][v];
    }
    assign_codes(code[t], len[t], mn, mx, alpha);
  }
  *n_groups_out = n_groups;
  return n_sel;
}

/* src/compression/compress.cu:119-137 */
void orc_selector_mtf(const uint8_t *sel, int32_t n_sel, int32_t n_groups, uint8_t *out) {
  uint8_t pos[ORC_MAX_GROUPS];
  for (int i = 0; i < n_groups; i++) pos[i] = (uint8_t)i;
  for (int32_t i = 0; i < n_sel; i++) {
    uint8_t v = sel[i];
    int j = 0;
    uint8_t tmp = pos[0];
    while (v != tmp) {
      j++;
      uint8_t t2 = tmp;
      tmp = pos[j];
      pos[j] = t2;
    }
    pos[0] = tmp;
    out[i] = (uint8_t)j;
  }
}

/* src/compression/huffman_encode.cu:15-71: code of sym[i] from table
 * sel[i/50], bit offsets by exclusive scan, packed MSB-first into u32. */
int64_t orc_huff_encode(const uint16_t *sym, int32_t s,
                        uint8_t len[ORC_MAX_GROUPS][ORC_MAX_ALPHA],
                        int32_t code[ORC_MAX_GROUPS][ORC_MAX_ALPHA],
                        const uint8_t *sel, uint32_t *words) {
  int64_t bits = 0;
  for (int32_t i = 0; i < s; i++) bits += len[sel[i / ORC_G_SIZE]][sym[i]];
  int64_t n_words = (bits + 31) / 32;
  memset(words, 0, sizeof(uint32_t) * (size_t)(n_words + 1));
  int64_t off = 0;
  for (int32_t i = 0; i < s; i++) {
    int t = sel[i / ORC_G_SIZE];
    uint32_t l = len[t][sym[i]];
    uint32_t c = (uint32_t)code[t][sym[i]];
    int64_t w = off >> 5;
    uint32_t sh = (uint32_t)(off & 31);
    if (sh + l <= 32) {
      words[w] |= c << (32 - sh - l);
    } else {
      uint32_t second = sh + l - 32;
      words[w] |= c >> second;
      words[w + 1] |= c << (32 - second);
    }
    off += l;
  }
  return bits;
}

/* ---------------------------------------------------------- block driver */
/* src/compression/compress.cu:60-182 */
int orc_compress_block(const uint8_t *in, int32_t n, orc_block *b) {
  memset(b, 0, sizeof *b);
  if (n <= 0) return -1;
  b->n = n;
  b->crc = orc_crc32(in, (size_t)n);
  uint8_t *r1 = (uint8_t *)malloc((size_t)n * 5 / 4 + 16);
  int32_t m = orc_rle1(in, n, r1);
  b->m = m;
  int32_t *sa = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
  b->rounds = orc_bwt(r1, m, sa);
  uint8_t *mtf = (uint8_t *)malloc((size_t)m);
  b->n_in_use = orc_mtf(r1, sa, m, mtf, &b->orig_ptr, b->present);
  b->alpha_size = b->n_in_use + 2;
  uint16_t *sym = (uint16_t *)malloc(sizeof(uint16_t) * ((size_t)m + 2));
  b->s = orc_rle2(mtf, m, b->n_in_use, sym);
  int32_t n_sel_cap = (b->s + ORC_G_SIZE - 1) / ORC_G_SIZE;
  b->sel = (uint8_t *)malloc((size_t)n_sel_cap + 1);
  b->sel_mtf = (uint8_t *)malloc((size_t)n_sel_cap + 1);
  b->n_sel = orc_huff_build(sym, b->s, b->n_in_use, b->len, b->code, b->sel, &b->n_groups);
  orc_selector_mtf(b->sel, b->n_sel, b->n_groups, b->sel_mtf);
  b->words = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)b->s * 17 / 32 + 4));
  b->bits = orc_huff_encode(sym, b->s, b->len, b->code, b->sel, b->words);
  free(r1); free(sa); free(mtf); free(sym);
  return 0;
}

void orc_block_free(orc_block *b) {
  free(b->sel); free(b->sel_mtf); free(b->words);
  b->sel = b->sel_mtf = NULL;
  b->words = NULL;
}

/* ------------------------------------------------------------ bit writer */
/* src/compression/compress.cu:26-58: MSB-first bit writer. */
typedef struct {
  uint8_t *out;
  size_t cap, pos;
  uint64_t buf;
  int nbits;
  int overflow;
} bitw;

static void bw_put(bitw *w, uint32_t v, int n) {
  if (n == 0) return;
  v &= (uint32_t)((1ull << n) - 1);
  w->buf = (w->buf << n) | v;
  w->nbits += n;
  while (w->nbits >= 8) {
    w->nbits -= 8;
    if (w->pos < w->cap) w->out[w->pos] = (uint8_t)(w->buf >> w->nbits);
    else w->overflow = 1;
    w->pos++;
  }
}

static void bw_pad(bitw *w) {
  if (w->nbits > 0) bw_put(w, 0, 8 - w->nbits);
}

/* src/compression/compress.cu:200-204 and :284-375 */
int64_t orc_write_stream(const orc_block *blocks, int32_t n_blocks, int level,
                         uint8_t *out, size_t cap) {
  if (level < 1) level = 1;
  if (level > 9) level = 9;
  bitw w = {out, cap, 0, 0, 0, 0};
  bw_put(&w, 0x42, 8); bw_put(&w, 0x5A, 8); bw_put(&w, 0x68, 8);
  bw_put(&w, (uint32_t)(0x30 + level), 8);
  uint32_t combined = 0;
  for (int32_t i = 0; i < n_blocks; i++) {
    const orc_block *b = &blocks[i];
    combined = (combined << 1) | (combined >> 31);
    combined ^= b->crc;
    bw_put(&w, 0x31, 8); bw_put(&w, 0x41, 8); bw_put(&w, 0x59, 8);
    bw_put(&w, 0x26, 8); bw_put(&w, 0x53, 8); bw_put(&w, 0x59, 8);
    bw_put(&w, b->crc, 32);
    bw_put(&w, 0, 1);
    bw_put(&w, (uint32_t)b->orig_ptr, 24);
    uint32_t l1 = 0;
    for (int r = 0; r < 16; r++) {
      int any = 0;
      for (int q = 0; q < 16; q++) any |= b->present[r * 16 + q];
      if (any) l1 |= 1u << (15 - r);
    }
    bw_put(&w, l1, 16);
    for (int r = 0; r < 16; r++) {
      if (!((l1 >> (15 - r)) & 1)) continue;
      uint32_t l2 = 0;
      for (int q = 0; q < 16; q++)
        if (b->present[r * 16 + q]) l2 |= 1u << (15 - q);
      bw_put(&w, l2, 16);
    }
    bw_put(&w, (uint32_t)b->n_groups, 3);
    bw_put(&w, (uint32_t)b->n_sel, 15);
    for (int32_t k = 0; k < b->n_sel; k++) {
      for (int q = 0; q < b->sel_mtf[k]; q++) bw_put(&w, 1, 1);
      bw_put(&w, 0, 1);
    }
    for (int t = 0; t < b->n_groups; t++) {
      int32_t cur = b->len[t][0];
      bw_put(&w, (uint32_t)cur, 5);
      for (int32_t v = 0; v < b->alpha_size; v++) {
        int32_t target = b->len[t][v];
        while (cur < target) { bw_put(&w, 2, 2); cur++; }
        while (cur > target) { bw_put(&w, 3, 2); cur--; }
        bw_put(&w, 0, 1);
      }
    }
    int64_t rem = b->bits;
    int64_t nw = (b->bits + 31) / 32;
    for (int64_t k = 0; k < nw; k++) {
      int take = rem < 32 ? (int)rem : 32;
      bw_put(&w, b->words[k] >> (32 - take), take);
      rem -= take;
    }
  }
  bw_put(&w, 0x17, 8); bw_put(&w, 0x72, 8); bw_put(&w, 0x45, 8);
  bw_put(&w, 0x38, 8); bw_put(&w, 0x50, 8); bw_put(&w, 0x90, 8);
  bw_put(&w, combined, 32);
  bw_pad(&w);
  return w.overflow ? -1 : (int64_t)w.pos;
}

size_t orc_compress_bound(size_t len) {
  size_t blocks = len / 100000 + 2;
  return len * 3 + blocks * 16384 + 64;
}

typedef struct {
  const uint8_t *in;
  size_t len;
  int32_t block_bytes;
  int32_t n_blocks;
  orc_block *blocks;
  int next;
  pthread_mutex_t mu;
} mt_ctx;

static void *mt_worker(void *arg) {
  mt_ctx *c = (mt_ctx *)arg;
  for (;;) {
    pthread_mutex_lock(&c->mu);
    int i = c->next++;
    pthread_mutex_unlock(&c->mu);
    if (i >= c->n_blocks) break;
    size_t off = (size_t)i * (size_t)c->block_bytes;
    size_t n = c->len - off < (size_t)c->block_bytes ? c->len - off : (size_t)c->block_bytes;
    orc_compress_block(c->in + off, (int32_t)n, &c->blocks[i]);
  }
  return NULL;
}

/* src/compression/compress.cu:184-375 with the block cut of
 * src/utils/io.cpp:2-32 (fixed-size chunks; no empty block is emitted). */
int64_t orc_compress_mt(const uint8_t *in, size_t len, int level,
                        int32_t block_bytes, uint8_t *out, size_t cap, int threads) {
  if (block_bytes <= 0) return -1;
  int32_t n_blocks = (int32_t)((len + (size_t)block_bytes - 1) / (size_t)block_bytes);
  orc_block *blocks = (orc_block *)calloc((size_t)n_blocks + 1, sizeof(orc_block));
  mt_ctx c;
  c.in = in; c.len = len; c.block_bytes = block_bytes; c.n_blocks = n_blocks;
  c.blocks = blocks; c.next = 0;
  pthread_mutex_init(&c.mu, NULL);
  if (threads < 1) threads = 1;
  if (threads == 1 || n_blocks <= 1) {
    mt_worker(&c);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, mt_worker, &c);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
  }
  pthread_mutex_destroy(&c.mu);
  int64_t r = orc_write_stream(blocks, n_blocks, level, out, cap);
  for (int32_t i = 0; i < n_blocks; i++) orc_block_free(&blocks[i]);
  free(blocks);
  return r;
}

int64_t orc_compress(const uint8_t *in, size_t len, int level, int32_t block_bytes,
                     uint8_t *out, size_t cap) {
  return orc_compress_mt(in, len, level, block_bytes, out, cap, 1);
}
