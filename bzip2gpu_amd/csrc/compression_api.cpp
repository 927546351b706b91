// compression_api.cpp - the reference-shaped C++ stage API (compression.h)
// over the C ABI in bzg.h.  Allocation and exit-on-error behaviour follow
// the reference (see compression.h).
#include "compression.h"

#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
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

// compress.cu:184-375 - appends one .bz2 stream to `out`.  The chunk cut IS
// the block cut (io.cpp:17-18): chunks are fetched in chunk order and copied
// into one of two device staging buffers as they arrive; every filled
// buffer is compressed as one segment of the stream (bzg_compress_segment)
// on a worker thread while the other buffer fills, so reading, host-to-
// device copies and compression overlap and host memory holds only the
// ring's chunks and the output.
namespace {

constexpr size_t kDefaultSegmentBytes = size_t(256) << 20;

struct Staging {
  uint8_t *d = nullptr;
  size_t len = 0;
  bool ready = false;  // filled, waiting for the worker
};

// MSB-first bits appended to a byte vector whose last byte already holds
// `used` bits (0 = byte-aligned).
struct TailBits {
  std::vector<uint8_t> &out;
  int used;
  void put(uint32_t v, int k) {
    for (int i = k - 1; i >= 0; i--) {
      if (used == 0) out.push_back(0);
      if ((v >> i) & 1u) out.back() |= (uint8_t)(0x80u >> used);
      used = (used + 1) & 7;
    }
  }
};

}  // namespace

void bzip2_gpu_compress_segmented(BZFileInputStream &in_stream, int n, std::vector<uint8_t> &out,
                                  size_t segment_bytes) {
  if (n < 1) n = 1;
  if (n > 9) n = 9;
  const int32_t bb = in_stream.buffer_size;
  if (bb <= 0 || bb > BZG_REF_BLOCK_BYTES) {
    fprintf(stderr, "bzip2_gpu_compress: chunk size must be 1..900000\n");
    exit(EXIT_FAILURE);
  }
  const size_t per = segment_bytes / (size_t)bb > 0 ? segment_bytes / (size_t)bb : 1;  // chunks per segment
  const size_t cap = per * (size_t)bb;
  out.push_back('B');
  out.push_back('Z');
  out.push_back('h');
  out.push_back((uint8_t)('0' + n));
  hipStream_t copy_s = nullptr, comp_s = nullptr;
  check_hip(hipStreamCreateWithFlags(&copy_s, hipStreamNonBlocking), "hipStreamCreate");
  check_hip(hipStreamCreateWithFlags(&comp_s, hipStreamNonBlocking), "hipStreamCreate");
  Staging buf[2];
  std::mutex mu;
  std::condition_variable cv;
  bool done = false;
  uint32_t crc = 0;
  int used = 0;  // bits of out.back() the stream already holds
  std::thread worker([&] {
    std::vector<uint8_t> seg;
    for (int k = 0;; k ^= 1) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return buf[k].ready || done; });
        if (!buf[k].ready) return;  // done, nothing left
      }
      seg.resize(bzg_compress_bound(buf[k].len, bb));
      size_t got = 0;
      int64_t bits = 0;
      check(bzg_compress_segment(buf[k].d, buf[k].len, bb, used, &crc, seg.data(), seg.size(), &got, &bits, comp_s),
            "bzip2_gpu_compress");
      // the segment's first `used` bits are zero: they merge into out.back()
      size_t from = 0;
      if (used > 0 && got > 0) {
        out.back() |= seg[0];
        from = 1;
      }
      out.insert(out.end(), seg.begin() + (std::ptrdiff_t)from, seg.begin() + (std::ptrdiff_t)got);
      used = (int)(bits & 7);
      {
        std::lock_guard<std::mutex> lk(mu);
        buf[k].ready = false;
      }
      cv.notify_all();
    }
  });
  bool eof = false;
  for (int k = 0; !eof; k ^= 1) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return !buf[k].ready; });
    }
    size_t len = 0;
    for (size_t c = 0; c < per; c++) {
      uint8_t *chunk;
      int idx;
      const int l = in_stream.fetch(chunk, idx);
      if (l < 0) {
        eof = true;
        break;
      }
      if (!buf[k].d) check_hip(hipMalloc((void **)&buf[k].d, cap), "hipMalloc");
      check_hip(hipMemcpyAsync(buf[k].d + len, chunk, (size_t)l, hipMemcpyHostToDevice, copy_s), "hipMemcpyAsync");
      check_hip(hipStreamSynchronize(copy_s), "hipStreamSynchronize");
      (void)hipHostFree(chunk);
      len += (size_t)l;
    }
    if (len > 0) {
      {
        std::lock_guard<std::mutex> lk(mu);
        buf[k].len = len;
        buf[k].ready = true;
      }
      cv.notify_all();
    }
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    done = true;
  }
  cv.notify_all();
  worker.join();
  for (auto &b : buf)
    if (b.d) (void)hipFree(b.d);
  (void)hipStreamDestroy(copy_s);
  (void)hipStreamDestroy(comp_s);
  if (in_stream.failed) {
    fprintf(stderr, "bzip2_gpu_compress: input could not be read\n");
    exit(EXIT_FAILURE);
  }
  // end-of-stream magic, combined CRC, zero padding (compress.cu:367-375)
  TailBits tb{out, used};
  tb.put(0x177245u, 24);
  tb.put(0x385090u, 24);
  tb.put(crc, 32);
}

void bzip2_gpu_compress(BZFileInputStream &in_stream, int n, std::vector<uint8_t> &out) {
  bzip2_gpu_compress_segmented(in_stream, n, out, kDefaultSegmentBytes);
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
