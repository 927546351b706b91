/*
 * datagen.c - seeded synthetic inputs for the benchmark workloads.
 *
 * Text: order-1 / order-2 Markov chains over the 95 printable ASCII bytes
 * plus '\n' (96 symbols); every context row is Zipf(s) over its own seeded
 * permutation of the symbols.  Uniform: i.i.d. bytes.  Runs: uniform byte
 * values with Geometric(1/mean) run lengths.  Mix: 1 MiB segments drawn
 * 50/25/25 from order-2 text (s=1.2) / uniform / runs (mean 64).  Periodic:
 * a seeded random base string of length p repeated, with a fraction of
 * positions XOR-ed by a nonzero byte.  PRNG: splitmix64.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s; } rng_t;

static uint64_t next64(rng_t *r) {
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unif(rng_t *r) { return (double)(next64(r) >> 11) * (1.0 / 9007199254740992.0); }
static uint32_t below(rng_t *r, uint32_t n) { return (uint32_t)((next64(r) >> 32) * (uint64_t)n >> 32); }

#define NSYM 96
static uint8_t sym_of(int i) { return i < 95 ? (uint8_t)(0x20 + i) : (uint8_t)'\n'; }

/* cdf[k] for ranks 0..95 of Zipf(s) */
static void zipf_cdf(double s, double *cdf) {
  double tot = 0;
  for (int k = 0; k < NSYM; k++) tot += pow(k + 1.0, -s);
  double acc = 0;
  for (int k = 0; k < NSYM; k++) {
    acc += pow(k + 1.0, -s) / tot;
    cdf[k] = acc;
  }
  cdf[NSYM - 1] = 1.0;
}

static int draw_rank(rng_t *r, const double *cdf) {
  double u = unif(r);
  int lo = 0, hi = NSYM - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cdf[mid] < u) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static void perm(rng_t *r, uint8_t *p) {
  for (int i = 0; i < NSYM; i++) p[i] = (uint8_t)i;
  for (int i = NSYM - 1; i > 0; i--) {
    int j = (int)below(r, (uint32_t)i + 1);
    uint8_t t = p[i]; p[i] = p[j]; p[j] = t;
  }
}

/* order-k Markov text, k in {1, 2} */
void dg_markov(uint8_t *out, size_t n, uint64_t seed, int order, double s) {
  rng_t r = {seed};
  int ctx = order == 1 ? NSYM : NSYM * NSYM;
  uint8_t *rows = (uint8_t *)malloc((size_t)ctx * NSYM);
  for (int c = 0; c < ctx; c++) perm(&r, rows + (size_t)c * NSYM);
  double cdf[NSYM];
  zipf_cdf(s, cdf);
  int a = (int)below(&r, NSYM), b = (int)below(&r, NSYM);
  for (size_t i = 0; i < n; i++) {
    int c = order == 1 ? b : a * NSYM + b;
    int nx = rows[(size_t)c * NSYM + draw_rank(&r, cdf)];
    out[i] = sym_of(nx);
    a = b;
    b = nx;
  }
  free(rows);
}

void dg_uniform(uint8_t *out, size_t n, uint64_t seed) {
  rng_t r = {seed};
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t v = next64(&r);
    memcpy(out + i, &v, 8);
  }
  for (; i < n; i++) out[i] = (uint8_t)next64(&r);
}

void dg_runs(uint8_t *out, size_t n, uint64_t seed, double mean) {
  rng_t r = {seed};
  double p = 1.0 / mean;
  size_t i = 0;
  while (i < n) {
    uint8_t v = (uint8_t)next64(&r);
    double u = unif(&r);
    size_t len = 1 + (size_t)floor(log(1.0 - u) / log(1.0 - p));
    for (size_t k = 0; k < len && i < n; k++) out[i++] = v;
  }
}

/* 1 MiB segments: 50% order-2 text (Zipf 1.2), 25% uniform, 25% runs(64). */
void dg_mix(uint8_t *out, size_t n, uint64_t seed) {
  rng_t r = {seed};
  const size_t SEG = 1u << 20;
  for (size_t off = 0, k = 0; off < n; off += SEG, k++) {
    size_t len = n - off < SEG ? n - off : SEG;
    double u = unif(&r);
    uint64_t sub = seed * 1000003ull + k * 7919ull + 1;
    if (u < 0.5) dg_markov(out + off, len, sub, 2, 1.2);
    else if (u < 0.75) dg_uniform(out + off, len, sub);
    else dg_runs(out + off, len, sub, 64.0);
  }
}

void dg_periodic(uint8_t *out, size_t n, uint64_t seed, size_t period, double mut_frac) {
  rng_t r = {seed};
  uint8_t *base = (uint8_t *)malloc(period ? period : 1);
  for (size_t i = 0; i < period; i++) base[i] = (uint8_t)next64(&r);
  for (size_t i = 0; i < n; i++) out[i] = base[i % period];
  size_t nm = (size_t)floor(mut_frac * (double)n);
  for (size_t k = 0; k < nm; k++) {
    size_t pos = (size_t)(unif(&r) * (double)n);
    if (pos >= n) pos = n - 1;
    out[pos] ^= (uint8_t)(1 + below(&r, 255));
  }
  free(base);
}
