// refapi_test.cpp - the reference's own test cases, run through the drop-in
// C++ API (include/compression.h, include/io.h) that the reference's callers
// bind to, with every output checked against the CPU oracle.
//
// Each case regenerates its input exactly as the reference test does (glibc
// srand(2137) / rand()) and names the reference test it mirrors.  Device
// outputs are released the way the reference's tests release them
// (hipFree for callee-allocated device buffers, delete[] for host tables,
// hipHostFree for ingestion chunks).  The oracle (oracle/oracle.h) is the
// checker; the reference's own CPU spec functions are restated where a case
// asserts against them.
//
// Usage: refapi_test [name-substring]   (prints "[ OK ]" / "[ FAIL ]" lines;
// exit status = number of failed cases); refapi_test --list prints the case
// names without touching the GPU.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <fstream>
#include <functional>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <streambuf>
#include <string>
#include <vector>

#include "bzg.h"
#include "compression.h"
#include "io.h"
#include "oracle.h"

namespace {

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define EXPECT(cond, ...)                                   \
  do {                                                      \
    if (!(cond)) {                                          \
      char msg_[512];                                       \
      snprintf(msg_, sizeof msg_, __VA_ARGS__);             \
      throw Failure(std::string(#cond) + ": " + msg_);      \
    }                                                       \
  } while (0)

void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Failure(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T *to_device(const std::vector<T> &h) {
  T *d = nullptr;
  hip_ok(hipMalloc((void **)&d, std::max<size_t>(1, h.size()) * sizeof(T)), "hipMalloc");
  if (!h.empty()) hip_ok(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "H2D");
  return d;
}

template <typename T>
std::vector<T> to_host(const T *d, size_t n) {
  std::vector<T> h(n);
  if (n) hip_ok(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "D2H");
  return h;
}

template <typename T>
size_t first_diff(const std::vector<T> &a, const std::vector<T> &b) {
  size_t n = std::min(a.size(), b.size());
  for (size_t i = 0; i < n; i++)
    if (a[i] != b[i]) return i;
  return a.size() == b.size() ? (size_t)-1 : n;
}

std::string tmp_path(const char *tag) {
  const char *dir = getenv("TMPDIR");
  return std::string(dir && *dir ? dir : "/tmp") + "/bzg_refapi_" + std::to_string(getpid()) + "_" + tag;
}

// ------------------------------------------------------------------ inputs
// tests/compression/rle1.cu:34-48: runs of rand() % R + 1 bytes of a value
// that advances by rand() % 100 + 1.
std::vector<uint8_t> rle1_input(int max_run) {
  srand(2137);
  const int n = 1000000;
  std::vector<uint8_t> in(n);
  uint8_t chr = 0;
  int reps = 0;
  for (int i = 0; i < n; i++) {
    if (!reps) {
      reps = rand() % max_run + 1;
      chr += rand() % 100 + 1;
    }
    in[i] = chr;
    reps--;
  }
  return in;
}

// tests/compression/rle2.cu:35-47: runs of rand() % R + 1 copies of a
// rand() % 256 byte.
std::vector<uint8_t> rle2_input(int max_run) {
  srand(2137);
  const int n = 1000000;
  std::vector<uint8_t> in(n);
  int reps = 0;
  for (int i = 0; i < n; i++) {
    if (!reps) {
      reps = rand() % max_run + 1;
      in[i] = (uint8_t)(rand() % 256);
    } else {
      in[i] = in[i - 1];
    }
    reps--;
  }
  return in;
}

int present_count(const std::vector<uint8_t> &in) {
  bool p[256] = {false};
  for (uint8_t c : in) p[c] = true;
  return (int)std::count(p, p + 256, true);
}

// The rule the reference's rle1 test spec states (tests/compression/rle1.cu:
// 8-32): a maximal run is cut every 259 bytes; a piece of r < 4 bytes is
// copied, a longer one becomes 4 copies and the count r - 4.
std::vector<uint8_t> rle1_rule(const std::vector<uint8_t> &in) {
  std::vector<uint8_t> out;
  size_t i = 0;
  while (i < in.size()) {
    size_t j = i;
    while (j < in.size() && in[j] == in[i] && j - i < 259) j++;
    const size_t r = j - i;
    if (r < 4) out.insert(out.end(), r, in[i]);
    else {
      out.insert(out.end(), 4, in[i]);
      out.push_back((uint8_t)(r - 4));
    }
    i = j;
  }
  return out;
}

// ------------------------------------------------------------------ cases
// tests/compression/rle1.cu:66-91
void rle1_case(int max_run) {
  const std::vector<uint8_t> in = rle1_input(max_run);
  std::vector<uint8_t> want(in.size() / 4 * 5 + 8);
  want.resize((size_t)orc_rle1(in.data(), (int32_t)in.size(), want.data()));
  EXPECT(want == rle1_rule(in), "oracle disagrees with the spec rule");
  uint8_t *d_in = to_device(in);
  uint8_t *d_out = nullptr;
  const int m = rle1_compress(d_in, (int)in.size(), d_out);
  hip_ok(hipFree(d_in), "hipFree");
  EXPECT(d_out != nullptr, "no output buffer");
  const std::vector<uint8_t> got = to_host(d_out, (size_t)m);
  hip_ok(hipFree(d_out), "hipFree");
  EXPECT((size_t)m == want.size(), "length %d vs %zu", m, want.size());
  EXPECT(got == want, "first difference at %zu", first_diff(got, want));
  if (max_run == 3) EXPECT(got == in, "runs below 4 must pass unchanged");  // rle1_incompressible
}

// tests/compression/fbwt.cu:26-74 (and simple_fbwt, :12-24: rotations by
// their bytes, ties by the smaller start)
void fbwt_case(int alphabet) {
  srand(2137);
  const int n = 5000;
  std::vector<uint8_t> in(n);
  for (int i = 0; i < n; i++) in[i] = (uint8_t)(rand() % alphabet);
  std::vector<int32_t> brute(n);
  std::iota(brute.begin(), brute.end(), 0);
  std::stable_sort(brute.begin(), brute.end(), [&](int a, int b) {
    for (int k = 0; k < n; k++) {
      const uint8_t x = in[(a + k) % n], y = in[(b + k) % n];
      if (x != y) return x < y;
    }
    return false;  // identical rotations keep ascending start order
  });
  std::vector<int32_t> want(n);
  orc_bwt(in.data(), n, want.data());
  EXPECT(want == brute, "oracle disagrees with the brute-force sort");
  uint8_t *d_in = to_device(in);
  int *d_sa = nullptr;
  fbwt(d_in, n, d_sa);
  hip_ok(hipFree(d_in), "hipFree");
  EXPECT(d_sa != nullptr, "no suffix array");
  const std::vector<int32_t> got = to_host(d_sa, (size_t)n);
  hip_ok(hipFree(d_sa), "hipFree");
  EXPECT(got == want, "first difference at %zu", first_diff(got, want));
}

// tests/compression/fmtf.cu:21-46
void symbols_table_case() {
  srand(2137);
  std::vector<uint8_t> table;
  for (int i = 0; i < 256; i++)
    if (rand() % 3 == 0) table.push_back((uint8_t)i);
  std::vector<uint8_t> data(100000);
  for (auto &c : data) c = table[(size_t)(rand() % (int)table.size())];
  uint8_t *d = to_device(data);
  uint8_t *got = nullptr;
  const int count = make_symbols_table(d, (int)data.size(), got);
  hip_ok(hipFree(d), "hipFree");
  EXPECT(count == (int)table.size(), "count %d vs %zu", count, table.size());
  for (int i = 0; i < count; i++) EXPECT(got[i] == table[(size_t)i], "entry %d", i);
  delete[] got;
}

// tests/compression/fmtf.cu:76-124: fmtf over the shifted identity SA
// (SA[i] = (i + 1) mod n, so L[i] = in[i]); the reference compares with
// the textbook MTF of the input (simple_fmtf, :48-74).
void fmtf_case(int alphabet) {
  srand(2137);
  const int n = 1000000;
  std::vector<uint8_t> in(n);
  for (int i = 0; i < n; i++) in[i] = (uint8_t)(rand() % alphabet);
  std::vector<int32_t> sa(n);
  for (int i = 0; i < n; i++) sa[i] = (i + 1) % n;
  std::vector<uint8_t> want(n), plain(n);
  int32_t want_ptr = -1;
  uint8_t want_present[256];
  const int want_nuse = orc_mtf(in.data(), sa.data(), n, want.data(), &want_ptr, want_present);
  orc_mtf_plain(in.data(), n, plain.data());
  EXPECT(want == plain, "oracle MTF disagrees with the textbook MTF");
  uint8_t *d_in = to_device(in);
  int *d_sa = to_device(sa);
  uint8_t *d_out = nullptr;
  int orig_ptr = -1;
  bool present[256];
  const int nuse = fmtf(d_in, d_sa, n, d_out, orig_ptr, present);
  hip_ok(hipFree(d_in), "hipFree");
  hip_ok(hipFree(d_sa), "hipFree");
  EXPECT(d_out != nullptr, "no output buffer");
  const std::vector<uint8_t> got = to_host(d_out, (size_t)n);
  hip_ok(hipFree(d_out), "hipFree");
  EXPECT(got == want, "first difference at %zu", first_diff(got, want));
  EXPECT(orig_ptr == want_ptr, "orig_ptr %d vs %d", orig_ptr, want_ptr);
  EXPECT(nuse == want_nuse, "n_in_use %d vs %d", nuse, want_nuse);
  for (int i = 0; i < 256; i++) EXPECT(present[i] == (want_present[i] != 0), "present[%d]", i);
}

// tests/compression/rle2.cu:80-143
void rle2_case(const std::vector<uint8_t> &in) {
  const int n = (int)in.size();
  const int table_size = present_count(in);
  std::vector<uint16_t> want((size_t)n + 1);
  want.resize((size_t)orc_rle2(in.data(), n, table_size, want.data()));
  uint8_t *d_in = to_device(in);
  uint16_t *d_out = nullptr;
  uint32_t *d_len = nullptr;
  hip_ok(hipMalloc((void **)&d_out, sizeof(uint16_t) * ((size_t)n + 1)), "hipMalloc");
  hip_ok(hipMalloc((void **)&d_len, sizeof(uint32_t)), "hipMalloc");
  rle2_compress(d_in, n, d_out, d_len, table_size);
  const uint32_t len = to_host(d_len, 1)[0];
  const std::vector<uint16_t> got = to_host(d_out, len);
  hip_ok(hipFree(d_in), "hipFree");
  hip_ok(hipFree(d_out), "hipFree");
  hip_ok(hipFree(d_len), "hipFree");
  EXPECT(len == want.size(), "length %u vs %zu", len, want.size());
  EXPECT(got == want, "first difference at %zu", first_diff(got, want));
  EXPECT(got.back() == table_size + 1, "end-of-block symbol");
}

std::vector<uint8_t> rle2_incompressible_input() {  // rle2.cu:100-104
  srand(2137);
  std::vector<uint8_t> in(1000000);
  for (auto &c : in) c = (uint8_t)((rand() % 255) + 1);
  return in;
}

// tests/compression/huffman.cu:9-38 calls the count_frequencies kernel; the
// C ABI's histogram entry point stands in for it.
void histogram_case() {
  srand(2137);
  const int n = 10000;
  std::vector<uint16_t> data(n);
  std::vector<int32_t> want(258, 0);
  for (auto &v : data) {
    v = (uint16_t)((rand() % 300) % 258);
    want[v]++;
  }
  uint16_t *d_data = to_device(data);
  int32_t *d_freq = to_device(std::vector<int32_t>(258, 0));
  EXPECT(bzg_histogram(d_data, n, d_freq, nullptr) == BZG_OK, "%s", bzg_last_error());
  const std::vector<int32_t> got = to_host(d_freq, 258);
  hip_ok(hipFree(d_data), "hipFree");
  hip_ok(hipFree(d_freq), "hipFree");
  EXPECT(got == want, "first difference at %zu", first_diff(got, want));
}

// tests/compression/huffman.cu:40-70: every alphabet size 1..255; the
// reference asserts only selectors <= 5, the oracle fixes lengths, codes
// and selectors exactly.
void tree_builder_case() {
  const int n = 10000;
  std::vector<uint16_t> data(n);
  uint16_t *d_data = nullptr;
  hip_ok(hipMalloc((void **)&d_data, n * sizeof(uint16_t)), "hipMalloc");
  srand(2137);
  for (int alpha = 1; alpha < 256; alpha++) {
    for (auto &v : data) v = (uint16_t)(rand() % alpha);
    hip_ok(hipMemcpy(d_data, data.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice), "H2D");
    uint8_t len[max_n_groups][max_alphabet_size];
    int32_t code[max_n_groups][max_alphabet_size];
    uint8_t *sel = nullptr;
    int n_groups = 0;
    const int n_sel = huffman_build_trees(d_data, n, alpha, len, code, sel, n_groups, 0);
    uint8_t wlen[ORC_MAX_GROUPS][ORC_MAX_ALPHA];
    int32_t wcode[ORC_MAX_GROUPS][ORC_MAX_ALPHA];
    std::vector<uint8_t> wsel((size_t)(n + 49) / 50);
    int32_t wng = 0;
    const int wn = orc_huff_build(data.data(), n, alpha, wlen, wcode, wsel.data(), &wng);
    EXPECT(n_sel == wn && n_groups == wng, "alphabet %d: n_sel %d/%d n_groups %d/%d", alpha, n_sel, wn, n_groups, wng);
    for (int i = 0; i < n_sel; i++) {
      EXPECT(sel[i] <= 5, "alphabet %d: selector %d = %d", alpha, i, sel[i]);
      EXPECT(sel[i] == wsel[(size_t)i], "alphabet %d: selector %d", alpha, i);
    }
    for (int g = 0; g < n_groups; g++)
      for (int v = 0; v < alpha + 2; v++) {
        EXPECT(len[g][v] == wlen[g][v], "alphabet %d: len[%d][%d]", alpha, g, v);
        EXPECT(code[g][v] == wcode[g][v], "alphabet %d: code[%d][%d]", alpha, g, v);
      }
    delete[] sel;
  }
  hip_ok(hipFree(d_data), "hipFree");
}

// tests/compression/huffman.cu:161-225: build, encode, decode back by
// walking the canonical codes, plus word-for-word equality with the oracle.
void encoding_case() {
  const int n = 10000, alpha = 15;
  std::vector<uint16_t> data(n);
  srand(2137);
  int shift = 0;
  for (int i = 0; i < n; i++) {
    if (i % 50 == 0) shift = rand() % 6;
    data[i] = (uint16_t)(rand() % (alpha / 2) + shift);
  }
  uint16_t *d_data = to_device(data);
  uint8_t len[max_n_groups][max_alphabet_size];
  int32_t code[max_n_groups][max_alphabet_size];
  uint8_t *sel = nullptr;
  int n_groups = 0;
  const int n_sel = huffman_build_trees(d_data, n, alpha, len, code, sel, n_groups, 0);
  uint8_t *d_sel = to_device(std::vector<uint8_t>(sel, sel + n_sel));
  uint32_t *d_words = nullptr;
  const int bits = huffman_encode(d_data, n, alpha, d_words, len, code, d_sel, n_sel, n_groups, 0);
  const std::vector<uint32_t> words = to_host(d_words, (size_t)(bits + 31) / 32);
  hip_ok(hipFree(d_data), "hipFree");
  hip_ok(hipFree(d_sel), "hipFree");
  hip_ok(hipFree(d_words), "hipFree");
  // decode: codes are prefix-free; match the next code of the group's table
  std::vector<uint16_t> dec;
  int64_t pos = 0;
  while (pos < bits && (int)dec.size() < n) {
    const int t = sel[dec.size() / 50];
    int found = -1;
    for (int v = 0; v < alpha + 2 && found < 0; v++) {
      const int l = len[t][v];
      if (l == 0 || pos + l > bits) continue;
      uint32_t x = 0;
      for (int k = 0; k < l; k++) {
        const int64_t b = pos + k;
        x = (x << 1) | ((words[(size_t)(b >> 5)] >> (31 - (b & 31))) & 1u);
      }
      if ((int32_t)x == code[t][v]) found = v;
      if (found >= 0) pos += l;
    }
    EXPECT(found >= 0, "no code matches at bit %lld", (long long)pos);
    dec.push_back((uint16_t)found);
  }
  EXPECT(dec == data, "decoded symbols differ at %zu", first_diff(dec, data));
  std::vector<uint32_t> want((size_t)n * 17 / 32 + 2);
  uint8_t wlen[ORC_MAX_GROUPS][ORC_MAX_ALPHA];
  int32_t wcode[ORC_MAX_GROUPS][ORC_MAX_ALPHA];
  memcpy(wlen, len, sizeof wlen);
  memcpy(wcode, code, sizeof wcode);
  const int64_t wbits = orc_huff_encode(data.data(), n, wlen, wcode, sel, want.data());
  EXPECT(wbits == bits, "bits %d vs %lld", bits, (long long)wbits);
  want.resize(words.size());
  EXPECT(words == want, "first word difference at %zu", first_diff(words, want));
  delete[] sel;
}

// Out-of-range packing input is refused by the C ABI (a code of 18 bits,
// a selector naming a table that does not exist).
void encode_rejects_case() {
  const int n = 100;
  std::vector<uint16_t> data(n, 1);
  uint16_t *d_data = to_device(data);
  std::vector<uint8_t> len(6 * 258, 2);
  std::vector<int32_t> code(6 * 258, 1);
  uint32_t *d_words = nullptr;
  hip_ok(hipMalloc((void **)&d_words, 4096), "hipMalloc");
  int64_t bits = 0;
  uint8_t *d_sel = to_device(std::vector<uint8_t>(2, 0));
  EXPECT(bzg_huff_encode(d_data, n, len.data(), code.data(), d_sel, 2, 2, d_words, &bits, nullptr) == BZG_OK, "%s",
         bzg_last_error());
  EXPECT(bits == 2 * n, "bits %lld", (long long)bits);
  len[258 + 1] = 18;
  EXPECT(bzg_huff_encode(d_data, n, len.data(), code.data(), d_sel, 2, 2, d_words, &bits, nullptr) == BZG_ERR_INVALID,
         "18-bit code accepted");
  len[258 + 1] = 2;
  hip_ok(hipFree(d_sel), "hipFree");
  d_sel = to_device(std::vector<uint8_t>{0, 2});
  EXPECT(bzg_huff_encode(d_data, n, len.data(), code.data(), d_sel, 2, 2, d_words, &bits, nullptr) == BZG_ERR_INVALID,
         "selector >= n_groups accepted");
  hip_ok(hipFree(d_sel), "hipFree");
  data[7] = 300;
  hip_ok(hipFree(d_data), "hipFree");
  d_data = to_device(data);
  d_sel = to_device(std::vector<uint8_t>(2, 0));
  EXPECT(bzg_huff_encode(d_data, n, len.data(), code.data(), d_sel, 2, 2, d_words, &bits, nullptr) == BZG_ERR_INVALID,
         "symbol >= 258 accepted");
  hip_ok(hipFree(d_sel), "hipFree");
  hip_ok(hipFree(d_data), "hipFree");
  hip_ok(hipFree(d_words), "hipFree");
}

// tests/io.cpp:6-26 with the reference's fixture (test_files/input_test_file,
// "word0\n" .. "word9\n") written afresh.
void file_input_stream_case() {
  const std::string path = tmp_path("io_fixture");
  {
    std::ofstream f(path, std::ios::binary);
    for (int i = 0; i < 10; i++) f << "word" << i << "\n";
  }
  int i = 0;
  {
    BZFileInputStream in(10, 6, path);
    uint8_t *buf;
    int len, chunk_idx = 0;
    while ((len = in.fetch(buf, chunk_idx)) > 0) {
      EXPECT(chunk_idx == i, "chunk %d arrived as %d", i, chunk_idx);
      const std::string want = "word" + std::to_string(i++) + "\n";
      EXPECT(len == (int)want.size() && memcmp(buf, want.data(), want.size()) == 0, "chunk %d content", i - 1);
      hip_ok(hipHostFree(buf), "hipHostFree");
    }
    EXPECT(!in.failed, "stream reported a failure");
  }
  unlink(path.c_str());
  EXPECT(i == 10, "%d chunks", i);
}

// A source that serves `good` bytes and then fails with a read error.
struct FailingBuf : std::streambuf {
  std::string data;
  bool served = false;
  explicit FailingBuf(size_t good) : data(good, 'x') {}
  int_type underflow() override {
    if (!served) {
      served = true;
      setg(&data[0], &data[0], &data[0] + data.size());
      return traits_type::to_int_type(data[0]);
    }
    throw std::ios_base::failure("device read error");
  }
};

// A read error part-way through the input ends the chunk stream with
// `failed` set, after the chunks read before it.  (The good bytes are a
// whole number of chunks: istream::read reports no count for a read that
// an exception cuts short.)
void read_error_case() {
  FailingBuf sb(24);
  std::istream is(&sb);
  BZFileInputStream in(4, 8, &is);
  uint8_t *buf;
  int len, idx, total = 0, chunks = 0;
  while ((len = in.fetch(buf, idx)) > 0) {
    total += len;
    chunks++;
    hip_ok(hipHostFree(buf), "hipHostFree");
  }
  EXPECT(in.failed, "read error not reported");
  EXPECT(total == 24 && chunks == 3, "%d bytes in %d chunks", total, chunks);
}

// Whole path through the ring: bzip2_gpu_compress on a file cut into the
// CLI's 900000-byte chunks, against the oracle's stream; the segmented
// variant with 2-chunk segments (so blocks land at every bit offset of a
// byte across segments) gives the same bytes.
std::vector<uint8_t> ring_input(size_t n, uint32_t seed) {
  std::vector<uint8_t> v(n);
  uint32_t x = seed;
  size_t i = 0;
  while (i < n) {
    x = x * 1664525u + 1013904223u;
    const uint32_t kind = x >> 30;
    x = x * 1664525u + 1013904223u;
    size_t run = 1 + ((x >> 8) % (kind == 0 ? 300u : 7u));
    x = x * 1664525u + 1013904223u;
    const uint8_t c = kind == 1 ? (uint8_t)(x >> 24) : (uint8_t)('a' + (x >> 24) % 26);
    for (size_t k = 0; k < run && i < n; k++) v[i++] = c;
  }
  return v;
}

void compress_case(size_t n, int level, size_t segment_bytes) {
  const std::vector<uint8_t> data = ring_input(n, (uint32_t)n + 7u);
  const std::string path = tmp_path("ring_input");
  {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(data.data()), (std::streamsize)data.size());
  }
  std::vector<uint8_t> want(orc_compress_bound(n));
  const int64_t wn = orc_compress(data.data(), n, level, 900000, want.data(), want.size());
  EXPECT(wn > 0, "oracle failed");
  want.resize((size_t)wn);
  std::vector<uint8_t> got{0xAB};  // bzip2_gpu_compress appends
  {
    BZFileInputStream in(10, 900000, path);
    if (segment_bytes) bzip2_gpu_compress_segmented(in, level, got, segment_bytes);
    else bzip2_gpu_compress(in, level, got);
  }
  unlink(path.c_str());
  EXPECT(got.size() == want.size() + 1 && got[0] == 0xAB, "size %zu vs %zu", got.size() - 1, want.size());
  got.erase(got.begin());
  EXPECT(got == want, "first difference at %zu", first_diff(got, want));
}

struct Case {
  const char *name;
  std::function<void()> fn;
};

}  // namespace

int main(int argc, char **argv) {
  const char *only = argc > 1 ? argv[1] : nullptr;
  const std::vector<Case> cases = {
      {"io.file_input_stream", file_input_stream_case},
      {"io.read_error_partway", read_error_case},
      {"compress.rle1_normal", [] { rle1_case(20); }},
      {"compress.rle1_incompressible", [] { rle1_case(3); }},
      {"compress.rle1_long_runs", [] { rle1_case(10000); }},
      {"compress.fbwt_normal", [] { fbwt_case(256); }},
      {"compress.fbwt_small_alphabet", [] { fbwt_case(4); }},
      {"compress.fbwt_single_character", [] { fbwt_case(1); }},
      {"compress.mtf_symbol_table_generation", symbols_table_case},
      {"compress.fmtf_normal", [] { fmtf_case(256); }},
      {"compress.fmtf_small_alphabet", [] { fmtf_case(4); }},
      {"compress.rle2_normal", [] { rle2_case(rle2_input(20)); }},
      {"compress.rle2_incompressible", [] { rle2_case(rle2_incompressible_input()); }},
      {"compress.rle2_long_runs", [] { rle2_case(rle2_input(10000)); }},
      {"compress.huffman_count_frequencies", histogram_case},
      {"compress.huffman_tree_builder_stability", tree_builder_case},
      {"compress.huffman_encoding", encoding_case},
      {"compress.huffman_encode_rejects_out_of_range", encode_rejects_case},
      {"compress.stream_empty", [] { compress_case(0, 9, 0); }},
      {"compress.stream_via_ring", [] { compress_case(4500000 + 12345, 9, 0); }},
      {"compress.stream_via_ring_segmented", [] { compress_case(4500000 + 12345, 3, 2 * 900000); }},
  };
  if (only && strcmp(only, "--list") == 0) {
    for (const Case &c : cases) printf("%s\n", c.name);
    return 0;
  }
  int failed = 0, ran = 0;
  for (const Case &c : cases) {
    if (only && !strstr(c.name, only)) continue;
    ran++;
    printf("[ RUN  ] %s\n", c.name);
    fflush(stdout);
    try {
      c.fn();
      printf("[ OK   ] %s\n", c.name);
    } catch (const std::exception &e) {
      failed++;
      printf("[ FAIL ] %s: %s\n", c.name, e.what());
    }
    fflush(stdout);
  }
  printf("%d of %d cases failed\n", failed, ran);
  return failed;
}
