/*
 * oracle.h - CPU restatement of the reference bzip2 block-compression path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libbzg.so) never links or calls it.
 *
 * Every function restates the semantics of the reference implementation
 * (SebastianBoteroLeonik/bzip2gpu) and cites the file:line it follows.
 * Integer-exact throughout.
 *
 * Pinning (what fixes each function independently of this restatement):
 *   orc_crc32         CRC-32/BZIP2 check value of "123456789" = 0xFC891918,
 *                     and the reference's own crc32.cpp built into oracle/_ref.
 *   orc_rle1          the reference test spec simple_rle1_compress
 *                     (tests/compression/rle1.cu:8-32) on its seeded inputs,
 *                     closed-form run sizes, and libbz2 stream equality.
 *   orc_bwt           brute-force rotation sort with index tie-break
 *                     (tests/compression/fbwt.cu:12-24) on its seeded inputs,
 *                     inverse-BWT round trip, libbz2 stream equality.
 *   orc_mtf           textbook MTF spec (tests/compression/fmtf.cu:48-74).
 *   orc_rle2          spec simple_rle2_compress (tests/compression/rle2.cu:9-33).
 *   orc_huff_build    libbz2 whole-stream byte equality on blocks where the
 *                     reference and stock bzip2 coincide; Kraft/maxLen
 *                     invariants; decode round trip.
 *   orc_huff_encode,  libbz2 byte equality + libbz2 decode round trip +
 *   orc_compress      the 14-byte empty-stream known answer.
 */
#ifndef BZG_ORACLE_H
#define BZG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_GROUPS 6
#define ORC_MAX_ALPHA 258
#define ORC_G_SIZE 50

/* Per-block result, the same fields the reference keeps in BlockData
 * (src/compression/compress.cu:13-24). */
typedef struct orc_block {
  uint32_t crc;            /* CRC-32/BZIP2 of the raw block bytes */
  int32_t n;               /* raw block length */
  int32_t m;               /* RLE1 length */
  int32_t s;               /* RLE2 symbol count (incl. EOB) */
  int32_t orig_ptr;        /* row of rotation 0 in the sorted order */
  int32_t n_in_use;        /* distinct bytes in the RLE1 output */
  int32_t alpha_size;      /* n_in_use + 2 */
  uint8_t present[256];    /* presence flags of RLE1 bytes */
  int32_t n_groups;        /* Huffman tables in use (2..6) */
  int32_t n_sel;           /* number of 50-symbol groups */
  uint8_t len[ORC_MAX_GROUPS][ORC_MAX_ALPHA];
  int32_t code[ORC_MAX_GROUPS][ORC_MAX_ALPHA];
  int64_t bits;            /* payload bits W */
  int32_t rounds;          /* prefix-doubling rounds the oracle executed */
  /* variable-size outputs, malloc'ed, freed by orc_block_free */
  uint8_t *sel;            /* selectors (table id per group) */
  uint8_t *sel_mtf;        /* MTF-coded selectors */
  uint32_t *words;         /* payload, MSB-first u32 words */
} orc_block;

uint32_t orc_crc32(const uint8_t *data, size_t len);
uint32_t orc_crc32_update(uint32_t crc_reg, const uint8_t *data, size_t len);

/* RLE1 (src/compression/rle1.cu:13-58).  out must hold 5*ceil(n/4)+5 bytes. */
int32_t orc_rle1(const uint8_t *in, int32_t n, uint8_t *out);

/* BWT suffix array of cyclic rotations (src/compression/fbwt.cu:13-95).
 * Identical rotations ordered by ascending start index.  Returns the number
 * of doubling rounds executed before every class was distinct. */
int32_t orc_bwt(const uint8_t *in, int32_t m, int32_t *sa);

/* Forward MTF over L[i] = in[(sa[i]-1) mod m] (src/compression/fmtf.cu:40-97,
 * :147-230).  Returns n_in_use; writes orig_ptr and present[256]. */
int32_t orc_mtf(const uint8_t *in, const int32_t *sa, int32_t m, uint8_t *out,
                int32_t *orig_ptr, uint8_t *present);

/* Plain MTF of a byte sequence with the sorted-present initial list
 * (tests/compression/fmtf.cu:48-74). */
void orc_mtf_plain(const uint8_t *in, int32_t m, uint8_t *out);

/* RLE2 (src/compression/rle2.cu:11-77).  out must hold m+1 u16. Returns s. */
int32_t orc_rle2(const uint8_t *mtf, int32_t m, int32_t n_in_use, uint16_t *out);

/* Huffman table build (src/compression/huffman.cu:256-364).
 * alphabet_size = n_in_use (the function adds 2, like the reference).
 * sel must hold ceil(s/50) bytes.  Returns n_sel. */
int32_t orc_huff_build(const uint16_t *sym, int32_t s, int32_t alphabet_size,
                       uint8_t len[ORC_MAX_GROUPS][ORC_MAX_ALPHA],
                       int32_t code[ORC_MAX_GROUPS][ORC_MAX_ALPHA],
                       uint8_t *sel, int32_t *n_groups);

/* libbz2-style code lengths (src/compression/huffman.cu:51-124). */
void orc_make_code_lengths(const int32_t *freq, uint8_t *len, int32_t alpha_size,
                           int32_t max_len);

/* Selector MTF (src/compression/compress.cu:119-137). */
void orc_selector_mtf(const uint8_t *sel, int32_t n_sel, int32_t n_groups,
                      uint8_t *out);

/* MSB-first packing (src/compression/huffman_encode.cu:15-143).
 * words must hold ceil(17*s/32)+1 words.  Returns W (bits). */
int64_t orc_huff_encode(const uint16_t *sym, int32_t s,
                        uint8_t len[ORC_MAX_GROUPS][ORC_MAX_ALPHA],
                        int32_t code[ORC_MAX_GROUPS][ORC_MAX_ALPHA],
                        const uint8_t *sel, uint32_t *words);

/* Whole per-block pipeline (src/compression/compress.cu:60-182).
 * n must be > 0.  Returns 0 on success. */
int orc_compress_block(const uint8_t *in, int32_t n, orc_block *out);
void orc_block_free(orc_block *b);

/* Stream writer (src/compression/compress.cu:184-375): header 'BZh'+level,
 * raw input cut into block_bytes chunks (the reference uses 900000 at every
 * level), empty chunks never emitted.  Returns bytes written, or -1 if cap
 * is too small.  Worst-case size: orc_compress_bound(len). */
int64_t orc_compress(const uint8_t *in, size_t len, int level,
                     int32_t block_bytes, uint8_t *out, size_t cap);
size_t orc_compress_bound(size_t len);

/* Block-parallel variant of orc_compress on `threads` host threads (the
 * reference's worker model, compress.cu:225-278); identical output. */
int64_t orc_compress_mt(const uint8_t *in, size_t len, int level,
                        int32_t block_bytes, uint8_t *out, size_t cap,
                        int threads);

/* Serialise already-computed blocks (compress.cu:284-375). */
int64_t orc_write_stream(const orc_block *blocks, int32_t n_blocks, int level,
                         uint8_t *out, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
