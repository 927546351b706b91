// compression.h - drop-in for the reference's stage API
// (SebastianBoteroLeonik/bzip2gpu include/compression.h:17-94), with
// hipStream_t in place of cudaStream_t.  Same names, argument meaning and
// ownership: functions that take `T *&out` allocate it stream-ordered with
// hipMallocAsync and the caller releases it with hipFreeAsync/hipFree;
// host arrays returned through `T *&` are new[]'d and the caller delete[]s
// them.  Any failure prints a message to stderr and exits with
// EXIT_FAILURE, as the reference's CUDA_ERROR_CHECK does (include/utils.h:3-12).
#ifndef BZG_COMPRESSION_H
#define BZG_COMPRESSION_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "io.h"

int rle1_compress(const uint8_t *in, int in_len, uint8_t *&out, hipStream_t stream = 0);

void rle2_compress(const uint8_t *in, int in_len, uint16_t *out, uint32_t *out_len,
                   int table_size, hipStream_t stream = 0);

int make_symbols_table(const uint8_t *d_in, int d_in_len, uint8_t *&symbols_table,
                       hipStream_t stream = 0);

int fmtf(const uint8_t *in_original, const int *in_suffix_array, int in_len, uint8_t *&out,
         int &orig_ptr, bool *present_symbols = nullptr, hipStream_t stream = 0);

constexpr int max_n_groups = 6;
constexpr int max_alphabet_size = 258;

int huffman_build_trees(uint16_t *device_data_in, int data_in_len, int alphabet_size,
                        uint8_t len[max_n_groups][max_alphabet_size],
                        int32_t code[max_n_groups][max_alphabet_size], uint8_t *&selectors,
                        int &n_groups, hipStream_t stream);

void fbwt(const uint8_t *in, int in_len, int *&out, hipStream_t stream = 0);

void bzip2_gpu_compress(BZFileInputStream &in_stream, int n, std::vector<uint8_t> &out);

int huffman_encode(uint16_t *dev_data_in, int data_in_len, int alphabet_size,
                   uint32_t *&dev_encoded_data, uint8_t len[max_n_groups][max_alphabet_size],
                   int32_t code[max_n_groups][max_alphabet_size], uint8_t *dev_selectors,
                   int32_t num_selectors, int n_groups, hipStream_t stream);

// Extension with no counterpart in the reference: bzip2_gpu_compress with an
// explicit segment size, the raw bytes compressed per pipeline run while
// the next segment is read (rounded down to whole chunks, at least one
// chunk; bzip2_gpu_compress uses 256 MiB).  The output does not depend on it.
void bzip2_gpu_compress_segmented(BZFileInputStream &in_stream, int n, std::vector<uint8_t> &out,
                                  size_t segment_bytes);

// Extension with no counterpart in the reference, whose decompression is not
// built (TODO.md:28-34): decodes the concatenated .bz2 streams in[0, len) on
// the device into `out` (replaced).  Returns 0; 2 for malformed or corrupt
// input (bad header, CRC mismatch, truncation); 3 for a device failure.  The
// message of a failure goes to *err when given.  Does not exit on errors.
int bzip2_gpu_decompress(const uint8_t *in, size_t len, std::vector<uint8_t> &out, std::string *err = nullptr);

#endif  // BZG_COMPRESSION_H
