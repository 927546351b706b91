// compression_api.cpp - the reference-shaped C++ stage API (compression.h)
// over the C ABI in bzg.h.  Allocation and exit-on-error behaviour follow
// the reference (see compression.h).
#include "compression.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "bzg.h"

namespace {

[[noreturn]] void die(const char *what) {
  fprintf(stderr, "%s failed: %s\n", what, bzg_last_error());
  exit(EXIT_FAILURE);
}

void check(bzg_status st, const char *what) {
  if (st != BZG_OK) die(what);
}

void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "%s failed! Error: %s\n\t %s\n", what, hipGetErrorName(e), hipGetErrorString(e));
    exit(EXIT_FAILURE);
  }
}

}  // namespace

// rle1.cu:60-113 - callee-allocated output (rle1.cu:99), empty input returns 0
int rle1_compress(const uint8_t *in, int in_len, uint8_t *&out, hipStream_t stream) {
  if (in_len == 0) return 0;
  uint8_t *tmp = nullptr;
  check_hip(hipMallocAsync((void **)&tmp, (size_t)bzg_rle1_bound(in_len), stream), "hipMallocAsync");
  int32_t m = 0;
  check(bzg_rle1(in, in_len, tmp, &m, stream), "rle1_compress");
  out = tmp;
  return m;
}

// rle2.cu:98-158 - caller-allocated output
void rle2_compress(const uint8_t *in, int in_len, uint16_t *out, uint32_t *out_len, int table_size,
                   hipStream_t stream) {
  check(bzg_rle2(in, in_len, table_size, out, out_len, stream), "rle2_compress");
}

// fmtf.cu:114-141 - host table from new[]
int make_symbols_table(const uint8_t *d_in, int d_in_len, uint8_t *&symbols_table, hipStream_t stream) {
  uint8_t table[256];
  int32_t count = 0;
  check(bzg_symbols_table(d_in, d_in_len, table, &count, stream), "make_symbols_table");
  symbols_table = new uint8_t[count > 0 ? count : 1];
  memcpy(symbols_table, table, (size_t)count);
  return count;
}

// fmtf.cu:147-230 - callee-allocated output (fmtf.cu:157-158)
int fmtf(const uint8_t *in_original, const int *in_suffix_array, int in_len, uint8_t *&out, int &orig_ptr,
         bool *present_symbols, hipStream_t stream) {
  if (in_len == 0) return 0;
  uint8_t *tmp = nullptr;
  check_hip(hipMallocAsync((void **)&tmp, (size_t)in_len, stream), "hipMallocAsync");
  int32_t op = 0, nuse = 0;
  uint8_t pres[256];
  check(bzg_mtf(in_original, in_suffix_array, in_len, tmp, &op, pres, &nuse, stream), "fmtf");
  if (present_symbols)
    for (int i = 0; i < 256; i++) present_symbols[i] = pres[i] != 0;
  out = tmp;
  orig_ptr = op;
  return nuse;
}

// huffman.cu:256-364 - host selectors from new[]
int huffman_build_trees(uint16_t *device_data_in, int data_in_len, int alphabet_size,
                        uint8_t len[max_n_groups][max_alphabet_size], int32_t code[max_n_groups][max_alphabet_size],
                        uint8_t *&selectors, int &n_groups, hipStream_t stream) {
  int32_t ng = 0, ns = 0;
  std::vector<uint8_t> sel((size_t)data_in_len / 50 + 2);
  check(bzg_huff_build(device_data_in, data_in_len, alphabet_size, &len[0][0], &code[0][0], sel.data(), &ng, &ns,
                       stream),
        "huffman_build_trees");
  selectors = new uint8_t[ns > 0 ? ns : 1];
  memcpy(selectors, sel.data(), (size_t)ns);
  n_groups = ng;
  return ns;
}

// fbwt.cu:47-105 - callee-allocated suffix array (fbwt.cu:51)
void fbwt(const uint8_t *in, int in_len, int *&out, hipStream_t stream) {
  if (in_len == 0) return;
  int *tmp = nullptr;
  check_hip(hipMallocAsync((void **)&tmp, (size_t)in_len * sizeof(int), stream), "hipMallocAsync");
  check(bzg_bwt(in, in_len, tmp, nullptr, stream), "fbwt");
  out = tmp;
}

// huffman_encode.cu:72-143 - callee-allocated words (huffman_encode.cu:124-125)
int huffman_encode(uint16_t *dev_data_in, int data_in_len, int /*alphabet_size*/, uint32_t *&dev_encoded_data,
                   uint8_t len[max_n_groups][max_alphabet_size], int32_t code[max_n_groups][max_alphabet_size],
                   uint8_t *dev_selectors, int32_t num_selectors, int n_groups, hipStream_t stream) {
  size_t words = (size_t)data_in_len * 17 / 32 + 4;
  uint32_t *tmp = nullptr;
  check_hip(hipMallocAsync((void **)&tmp, words * 4, stream), "hipMallocAsync");
  int64_t bits = 0;
  check(bzg_huff_encode(dev_data_in, data_in_len, &len[0][0], &code[0][0], dev_selectors, num_selectors, n_groups,
                        tmp, &bits, stream),
        "huffman_encode");
  dev_encoded_data = tmp;
  return (int)bits;
}

// compress.cu:184-375 - fetch every chunk in chunk order, then one batched
// pipeline call with the chunk size as block size (the chunk cut IS the
// block cut in the reference, io.cpp:17-18).  Appends to `out`.
void bzip2_gpu_compress(BZFileInputStream &in_stream, int n, std::vector<uint8_t> &out) {
  if (n < 1) n = 1;
  if (n > 9) n = 9;
  std::map<int, std::pair<uint8_t *, int>> chunks;
  uint8_t *buf;
  int idx;
  int len;
  while ((len = in_stream.fetch(buf, idx)) >= 0) chunks[idx] = {buf, len};
  if (in_stream.failed) {
    fprintf(stderr, "bzip2_gpu_compress: input could not be read\n");
    exit(EXIT_FAILURE);
  }
  std::vector<uint8_t> all;
  size_t total = 0;
  for (auto &kv : chunks) total += (size_t)kv.second.second;
  all.reserve(total);
  for (auto &kv : chunks) {
    all.insert(all.end(), kv.second.first, kv.second.first + kv.second.second);
    (void)hipHostFree(kv.second.first);
  }
  int32_t bb = in_stream.buffer_size;
  if (bb <= 0 || bb > BZG_REF_BLOCK_BYTES) {
    fprintf(stderr, "bzip2_gpu_compress: chunk size must be 1..900000\n");
    exit(EXIT_FAILURE);
  }
  size_t cap = bzg_compress_bound(total, bb);
  size_t base = out.size();
  out.resize(base + cap);
  size_t got = 0;
  check(bzg_compress_buffer(all.data(), total, n, bb, out.data() + base, cap, &got), "bzip2_gpu_compress");
  out.resize(base + got);
}

int bzip2_gpu_decompress(const uint8_t *in, size_t len, std::vector<uint8_t> &out, std::string *err) {
  size_t cap = len * 4 > (size_t)(1 << 20) ? len * 4 : (size_t)(1 << 20);
  out.resize(cap);
  size_t got = 0;
  bzg_status st = bzg_decompress_buffer(in, len, out.data(), out.size(), &got);
  if (st == BZG_ERR_CAPACITY) {
    out.resize(got);
    st = bzg_decompress_fetch(out.data(), out.size(), &got);
  }
  if (st != BZG_OK) {
    if (err) *err = bzg_last_error();
    out.clear();
    return st == BZG_ERR_DATA ? 2 : 3;
  }
  out.resize(got);
  return 0;
}
