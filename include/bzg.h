/*
 * bzg.h - C ABI of the MI355X bzip2 block-compression path.
 *
 * Plain C types only (pointers, sizes, status codes); streams are passed as
 * opaque `void*` (a hipStream_t, or NULL for the null stream).  Every entry
 * point returns a bzg_status; bzg_last_error() gives the message of the last
 * failure on the calling thread.  The HIP kernels behind these functions are
 * the only implementation: there is no CPU fallback.
 *
 * Each entry point names the reference interface it replaces
 * (SebastianBoteroLeonik/bzip2gpu).  Device pointers ("d_") must point to
 * device memory of the current device; host pointers ("h_") to host memory.
 */
#ifndef BZG_H
#define BZG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t bzg_status;
#define BZG_OK 0
#define BZG_ERR_INVALID 1  /* bad argument (size, level, null pointer) */
#define BZG_ERR_HIP 2      /* a HIP runtime call failed */
#define BZG_ERR_CAPACITY 3 /* output buffer too small */
#define BZG_ERR_INTERNAL 4 /* a device-side invariant failed */
#define BZG_ERR_DATA 5     /* malformed or corrupt .bz2 input (decompression) */

/* Raw bytes per block of the reference's stream driver: the input is cut
 * into fixed 900000-byte chunks at every level (src/main.cpp:26,36,
 * src/utils/io.cpp:17-18). */
#define BZG_REF_BLOCK_BYTES 900000

/* Version string of this library. */
const char *bzg_version(void);
/* Message of the last failure on this thread ("" if none). */
const char *bzg_last_error(void);
/* Selects and initialises a device (optional; device 0 by default). */
bzg_status bzg_init(int device);

/* ---------------------------------------------------------------- whole path
 * Replace bzip2_gpu_compress (include/compression.h:86-87,
 * src/compression/compress.cu:184-375): one .bz2 stream with header digit
 * `level` (clamped to 1..9) from `len` raw bytes cut into `block_bytes`
 * blocks (BZG_REF_BLOCK_BYTES reproduces the reference; 100000*level gives
 * stock-sized blocks).  Empty input yields the 14-byte empty stream.
 */
/* Worst-case size of the stream produced for `len` raw bytes cut into
 * `block_bytes` blocks. */
size_t bzg_compress_bound(size_t len, int32_t block_bytes);

/* Batched path: all blocks of the input in one device pipeline; stream
 * assembly on the device.  Host buffers in, host buffer out. */
bzg_status bzg_compress_buffer(const uint8_t *h_in, size_t len, int level,
                               int32_t block_bytes, uint8_t *h_out,
                               size_t out_cap, size_t *out_len);

/* Plain path: one block per pipeline launch on one stream, blocks
 * serialised on the host in order (the reference's compress_block +
 * BitWriter structure, compress.cu:26-182).  Same output bytes. */
bzg_status bzg_compress_plain(const uint8_t *h_in, size_t len, int level,
                              int32_t block_bytes, uint8_t *h_out,
                              size_t out_cap, size_t *out_len);

/* Plain path with the input already resident on the device (d_in, len
 * bytes); each block is compressed by its own pipeline run on `stream` and
 * its result copied back for host serialisation, as above. */
bzg_status bzg_compress_plain_device(const uint8_t *d_in, size_t len, int level,
                                     int32_t block_bytes, uint8_t *h_out,
                                     size_t out_cap, size_t *out_len,
                                     void *stream);

/* The reference's own concurrency for the plain path, over many inputs
 * already resident on the device (the worker pool of bzip2_gpu_compress,
 * src/compression/compress.cu:225-278, at most 16 threads): `threads` host
 * threads, each with its own pipeline and HIP stream, take the next block
 * of any input and compress it alone (one block per pipeline run, result
 * copied back); the host then serialises each input's stream in block
 * order.  Inputs as for bzg_compress_batch_device; stream i is written to
 * h_out + h_byte_off[i], h_bytes[i] bytes long (streams back to back).
 * Same bytes as the batched and single-stream plain paths. */
bzg_status bzg_compress_plain_threads(const uint8_t *d_in, int32_t n_inputs,
                                      const uint64_t *in_off,
                                      const uint64_t *lens,
                                      const int32_t *levels,
                                      int32_t block_bytes, int32_t threads,
                                      uint8_t *h_out, size_t out_cap,
                                      int64_t *h_byte_off, int64_t *h_bytes);

/* Batched path over many independent inputs already resident on the
 * device: input i is d_in[in_off[i] .. in_off[i]+lens[i]) and becomes its
 * own stream with level levels[i].  block_bytes <= 0 selects 100000*level
 * per input.  On return *d_out points to a device buffer owned by the
 * library (valid until the next call on this device) holding stream i at
 * h_byte_off[i], h_bytes[i] bytes long. */
bzg_status bzg_compress_batch_device(const uint8_t *d_in, int32_t n_inputs,
                                     const uint64_t *in_off,
                                     const uint64_t *lens,
                                     const int32_t *levels,
                                     int32_t block_bytes, uint8_t **d_out,
                                     int64_t *h_byte_off, int64_t *h_bytes,
                                     void *stream);

/* Allocation-free variant of bzg_compress_batch_device.  It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same inputs and stream semantics,
 * except that the streams go to the caller's d_out (out_cap bytes; stream i
 * at d_out + h_byte_off[i], h_bytes[i] bytes long) and the pipeline's
 * workspace is a fixed array of this library in device memory
 * (BZG_NOALLOC_SCRATCH_MIB MiB per GPU, set up when the library first runs
 * on the GPU and held until the process ends) instead of memory it
 * allocates.  It takes at most BZG_NOALLOC_MAX_BYTES raw bytes in total, cut
 * into at most BZG_NOALLOC_MAX_BLOCKS blocks; bigger batches give
 * BZG_ERR_INVALID, and an out_cap below the worst-case size of the planned
 * streams gives BZG_ERR_CAPACITY.  All calls on one GPU share its array,
 * which keeps the last call's contents (every run initialises what it
 * reads); the library serialises the calls and each returns after its work
 * on `stream` has finished, so two calls never overlap on the array, but
 * device work that did overlap a call (e.g. the same kernels replayed from a
 * capture) would race on it. */
#define BZG_NOALLOC_SCRATCH_MIB 96
#define BZG_NOALLOC_MAX_BYTES (1u << 20)
#define BZG_NOALLOC_MAX_BLOCKS 64
bzg_status bzg_compress_batch_device_noalloc(const uint8_t *d_in, int32_t n_inputs,
                                             const uint64_t *in_off,
                                             const uint64_t *lens,
                                             const int32_t *levels,
                                             int32_t block_bytes,
                                             uint8_t *d_out, size_t out_cap,
                                             int64_t *h_byte_off,
                                             int64_t *h_bytes, void *stream);

/* One segment of a stream compressed piece by piece, so that ingestion
 * overlaps compression (bzip2_gpu_compress's chunk loop over the reading
 * ring, src/compression/compress.cu:225-278 with src/utils/io.cpp:34-55,
 * and the BitWriter's block serialisation, compress.cu:284-365): the blocks
 * of d_in[0, len) (device memory), cut into block_bytes blocks, written as
 * bare block bits with no stream header and no end-of-stream trailer,
 * starting lead_bits (0..7) bits into the first output byte; those lead
 * bits are zero, for the caller to OR into the last, partial byte of what
 * it already holds.  *h_crc is the combined stream CRC: on entry its value
 * after the earlier segments (0 before the first), on return after this
 * one.  *h_bits receives lead_bits + this segment's block bits, so the
 * next segment's lead_bits is *h_bits mod 8.  The caller writes the stream
 * header ("BZh" + level) before the first segment and the 48-bit
 * end-of-stream magic, the final combined CRC and the zero padding after
 * the last.  Output size is bounded by bzg_compress_bound(len, block_bytes). */
bzg_status bzg_compress_segment(const uint8_t *d_in, size_t len,
                                int32_t block_bytes, int32_t lead_bits,
                                uint32_t *h_crc, uint8_t *h_out,
                                size_t out_cap, size_t *out_len,
                                int64_t *h_bits, void *stream);

/* Largest raw byte count one device pipeline run takes in the batched
 * entry points (default 1 GiB); bigger input sets run as consecutive
 * sub-batches with bounded device memory (about 70 bytes per input byte). */
bzg_status bzg_set_max_batch_bytes(uint64_t bytes);

/* Overlap of consecutive sub-batches (default on; no counterpart in the
 * reference): with two or more sub-batches, they alternate between two
 * pipelines on two HIP streams driven by two host threads, and sub-batch
 * p+1's CRC/RLE1 and suffix sort run while sub-batch p's MTF, Huffman,
 * packing and assembly do (the sorts stay in sub-batch order).  Calls that
 * time autotuner variants run their sub-batches one at a time.  Output is
 * byte-identical either way; the second pipeline holds its own workspace. */
bzg_status bzg_set_batch_overlap(int32_t on);

/* Per-stage times of the last batched call on this thread's device (ms,
 * HIP events), the number of suffix-sort rounds, and the suffix-sort rows
 * keyed and sorted, summed over rounds. */
bzg_status bzg_last_stage_times(float *h_ms6, int32_t *h_rounds,
                                int64_t *h_bwt_rows);

/* Stage graphs (no counterpart in the reference): the batched path runs the
 * CRC/RLE1, MTF/RLE2, Huffman, pack and assembly stage groups and the
 * suffix-sort rounds (round 0, and the later round replayed until the group
 * list runs empty) as HIP graphs captured once per device-side plan and
 * replayed while that plan recurs.  Every suffix-sort kernel takes its sizes
 * from device memory, so no round needs a host readback; the host checks
 * for the end once per chunk of rounds.  bzg_set_stage_graphs(0) runs
 * everything eagerly (also BZG_NO_GRAPHS=1 at start-up), as does an active
 * kernel probe (bzg_probe_kernel).  bzg_graph_captures / bzg_graph_replays
 * count captures and graph launches of this thread's device context (-1 on
 * error). */
bzg_status bzg_set_stage_graphs(int32_t on);
int64_t bzg_graph_captures(void);
int64_t bzg_graph_replays(void);

/* Launch-shape autotuner (no counterpart in the reference).  Every variant
 * of every knob gives identical output.
 * Knob 0 is the threads per workgroup of the large-group radix scatter,
 * variants 0/1/2 = 256/512/1024, chosen per size class (0: suffix-sort
 * round 0, whole blocks; 1: the later rounds) from the radix passes' time
 * per row and pass (rounds run one at a time while it is being timed).
 * Knob 1 is the launch policy of the suffix-sort rounds, variants 0/1/2 =
 * per round (a readback of the round's totals, then one workgroup per work
 * item) / device-driven (rounds plan themselves on the device and are
 * enqueued or replayed as graphs in chunks; round 0 on grids bounded by the
 * plan, later rounds on persistent grids of one or four residencies of the
 * device), chosen per batch size class
 * (0: under 8 MiB of block capacity, 1: under 128 MiB, 2: larger) from the
 * sort stage's time per sorted row, once knob 0 has settled.
 * Knob 2 is the number of consecutive 4096-symbol MTF segments one wave
 * takes, variants 0/1/2/3 = 1/2/4/8 (the list state entering the first
 * segment of a run is computed from the segment summaries, later segments
 * continue from the previous one), per size class (0: under 32 MiB of
 * block capacity, 1: larger), chosen from the MTF + RLE2 stage's time per
 * capacity tile.
 * Batched calls take their sub-batches one at a time while any knob is
 * being timed.
 * A choice is made after timing each variant twice (the least total time
 * per total work of its samples) and is kept in the file named by
 * BZG_TUNE_CACHE when that variable is set.
 * bzg_tune_force pins a variant for every class (-1: automatic).
 * bzg_tune_state returns the chosen variant of a class (-1 while exploring,
 * -2 for a bad argument) and fills, per variant, its launch value, the
 * nanoseconds per unit of work over its samples (knob 0: rows x radix
 * passes; knob 1: sorted rows; knob 2: capacity tiles), and its sample
 * count. */
bzg_status bzg_tune_force(int32_t knob, int32_t variant);
int32_t bzg_tune_state(int32_t knob, int32_t size_class, int32_t *h_values,
                       float *h_ns_per_unit, int32_t *h_samples,
                       int32_t max_variants);

/* Suffix-sort rows of the last batched call, summed over rounds: rows in
 * small groups (LDS sorts), rows in large groups (tile radix sort), and
 * large-group rows times the radix passes they went through (benchmarking
 * aid, no counterpart in the reference). */
bzg_status bzg_last_sort_rows(int64_t *h_small_rows, int64_t *h_large_rows,
                              int64_t *h_large_pass_rows);

/* Totals over the blocks of the last bzg_compress_batch_device call
 * (benchmarking aid, no counterpart in the reference): [0] blocks, [1] raw
 * bytes n, [2] RLE1 bytes m, [3] RLE2 symbols s incl. end-of-block, [4]
 * payload bits W, [5] the sum over blocks of m x suffix-sort rounds. */
bzg_status bzg_last_batch_totals(int64_t *h_totals6);

/* Rows sorted and GPU milliseconds of every prefix-doubling round of the
 * last pipeline run (its last sub-batch); returns the number of rounds and
 * fills at most max_rounds entries (either array may be NULL). */
int32_t bzg_last_round_stats(int64_t *h_rows, float *h_ms, int32_t max_rounds);

/* Kernel timing probe (no counterpart in the reference; for benchmarks):
 * bzg_probe_kernel(name) makes every later launch of the named kernel be
 * bracketed by HIP events on its stream ("" or NULL turns it off);
 * bzg_probe_read returns the summed duration and launch count since the
 * last read. */
bzg_status bzg_probe_kernel(const char *kernel_name);
bzg_status bzg_probe_read(float *h_total_ms, int32_t *h_launches);

/* ---------------------------------------------------------- decompression
 * The reference lists decompression (iBWT, iMTF, RLE2, Huffman decoding,
 * block-boundary discovery) as not built (TODO.md:28-34, SOURCES.md:5-6);
 * these entry points decode the format its stream driver writes
 * (compress.cu:184-375) and any stock bzip2 stream on the device.  Input is
 * one or more concatenated .bz2 streams; bytes after the last stream that
 * do not start a new "BZh" header are ignored (as Python's bz2 does).  Block
 * and stream CRCs are checked; a mismatch or malformed input returns
 * BZG_ERR_DATA.  Blocks may decode to up to 1125000 RLE1 bytes whatever the
 * level digit says (the reference's fixed 900000-byte cut can exceed
 * 100000*level); randomised blocks (bzip2 < 0.9.5) are rejected.
 */
/* Host buffers in and out.  If the output does not fit, returns
 * BZG_ERR_CAPACITY with *out_len set to the decoded size; the decoded bytes
 * stay on the device until the next decompression call on this device and
 * bzg_decompress_fetch copies them out. */
bzg_status bzg_decompress_buffer(const uint8_t *h_in, size_t len, uint8_t *h_out,
                                 size_t out_cap, size_t *out_len);
bzg_status bzg_decompress_fetch(uint8_t *h_out, size_t out_cap, size_t *out_len);

/* Input already on the device.  *d_out points to a device buffer owned by
 * the library (valid until the next decompression call on this device). */
bzg_status bzg_decompress_device(const uint8_t *d_in, size_t len, uint8_t **d_out,
                                 size_t *out_len, void *stream);

/* Stage times of the last decompression on this device (ms, HIP events):
 * [0] magic scan + Huffman decoding (incl. the block-chain walk on the host),
 * [1] inverse RLE2/MTF, [2] FL-map counting sort, [3] inverse BWT walks,
 * [4] inverse RLE1 + CRC; and the number of blocks and streams decoded. */
bzg_status bzg_last_decode_times(float *h_ms5, int32_t *h_blocks, int32_t *h_streams);

/* ---------------------------------------------------------------- serving
 * A request loop with continuous batching (no counterpart in the reference,
 * which compresses one input at a time).  Requests are device buffers; a
 * pool of `workers` pipelines, each on its own HIP stream, takes every
 * queued request that fits its batch (at most max_batch_bytes raw bytes,
 * and the device memory of all in-flight batches within hbm_budget_bytes)
 * and compresses them in one batched pipeline run, so batches of different
 * workers overlap on the device.  mode 1 serves each request alone on the
 * plain per-block path instead.  Block size is 100000*level per request.
 */
typedef struct bzg_server bzg_server;
bzg_status bzg_server_create(int32_t mode, int32_t workers, uint64_t max_batch_bytes,
                             uint64_t hbm_budget_bytes, bzg_server **out);
void bzg_server_destroy(bzg_server *srv);
/* Queues one request; *h_ticket identifies it (tickets increase
 * monotonically over the server's life and are never reused). */
bzg_status bzg_server_submit(bzg_server *srv, const uint8_t *d_in, uint64_t len, int32_t level,
                             int64_t *h_ticket);
/* Waits for a request; its stream is d_arena[*h_out_off .. +*h_out_bytes)
 * when outputs are kept (bzg_server_keep), else only the size is set.  The
 * request's record is released by the wait: each ticket is waited for once,
 * and a ticket that is unknown or already waited for gives BZG_ERR_INVALID. */
bzg_status bzg_server_wait(bzg_server *srv, int64_t ticket, int64_t *h_out_bytes, int64_t *h_out_off,
                           int64_t *h_latency_ns);
/* Keep outputs in the caller's device arena d_arena (arena_bytes), filled
 * from its start in completion order; outputs that no longer fit are not
 * kept (offset -1).  Waits until no request is queued or running first. */
bzg_status bzg_server_keep(bzg_server *srv, uint8_t *d_arena, uint64_t arena_bytes);
/* Open-loop replay: request i = d_pool[in_off[i] .. +lens[i]) at level
 * levels[i] is submitted at arrival_ns[i] (nanoseconds from the start,
 * non-decreasing); returns when all are done.  Fills completion times,
 * output sizes, arena offsets (-1 when not kept) and the size of the
 * pipeline run that served each request.  The replay first waits until
 * every request submitted earlier has finished (their tickets stay valid),
 * then restarts the server clock.  Calls of bzg_server_submit from other
 * threads during a replay join its mix. */
bzg_status bzg_server_replay(bzg_server *srv, const uint8_t *d_pool, int32_t n, const uint64_t *in_off,
                             const uint64_t *lens, const int32_t *levels, const int64_t *arrival_ns,
                             int64_t *h_done_ns, int64_t *h_out_bytes, int64_t *h_out_off,
                             int32_t *h_batch);

/* ------------------------------------------------------------- stage level
 * One block at a time, device buffers in and out, like the reference's
 * stage functions (include/compression.h:17-94).  Outputs are written to
 * caller-provided device buffers; the callee allocates nothing the caller
 * must free.
 */

/* bzip2_crc32 (src/compression/crc32.h:7, crc32.cpp:22-29). */
bzg_status bzg_crc32(const uint8_t *d_in, int32_t n, uint32_t *h_crc,
                     void *stream);

/* Stage entry points accept blocks of up to 16 MiB (1 << 24 elements).
 * Capacity d_out needs for bzg_rle1 on n raw bytes (>= 5*ceil(n/4)). */
int32_t bzg_rle1_bound(int32_t n);
/* rle1_compress (include/compression.h:17-18, rle1.cu:60-113). */
bzg_status bzg_rle1(const uint8_t *d_in, int32_t n, uint8_t *d_out,
                    int32_t *h_out_len, void *stream);

/* fbwt (include/compression.h:75, fbwt.cu:47-105): d_sa[m] = sorted
 * cyclic rotations, identical rotations in ascending index order.
 * *h_rounds = prefix-doubling rounds executed (may be NULL). */
bzg_status bzg_bwt(const uint8_t *d_in, int32_t m, int32_t *d_sa,
                   int32_t *h_rounds, void *stream);

/* make_symbols_table (include/compression.h:44-45, fmtf.cu:114-141):
 * h_table[256] receives the present bytes ascending, *h_count their number. */
bzg_status bzg_symbols_table(const uint8_t *d_in, int32_t n, uint8_t *h_table,
                             int32_t *h_count, void *stream);

/* fmtf (include/compression.h:54-56, fmtf.cu:147-230). h_present[256]
 * may be NULL.  Returns n_in_use in *h_n_in_use. */
bzg_status bzg_mtf(const uint8_t *d_rle1, const int32_t *d_sa, int32_t m,
                   uint8_t *d_out, int32_t *h_orig_ptr, uint8_t *h_present,
                   int32_t *h_n_in_use, void *stream);

/* rle2_compress (include/compression.h:31-32, rle2.cu:98-158): d_out needs
 * m+1 entries; *d_out_len (device) receives the symbol count incl. EOB. */
bzg_status bzg_rle2(const uint8_t *d_mtf, int32_t m, int32_t table_size,
                    uint16_t *d_out, uint32_t *d_out_len, void *stream);

/* count_frequencies (huffman.cu:126-133, called directly by
 * tests/compression/huffman.cu:7,29): d_freq[258] += histogram. */
bzg_status bzg_histogram(const uint16_t *d_sym, int32_t s, int32_t *d_freq,
                         void *stream);

/* huffman_build_trees (include/compression.h:67-72, huffman.cu:256-364):
 * alphabet_size = n_in_use (2 is added inside, as in the reference).
 * h_len[6*258], h_code[6*258] row-major; h_sel holds ceil(s/50) bytes. */
bzg_status bzg_huff_build(const uint16_t *d_sym, int32_t s,
                          int32_t alphabet_size, uint8_t *h_len,
                          int32_t *h_code, uint8_t *h_sel,
                          int32_t *h_n_groups, int32_t *h_n_sel,
                          void *stream);

/* huffman_encode (include/compression.h:89-94, huffman_encode.cu:72-143):
 * MSB-first packing into d_words (ceil(17*s/32)+2 words); *h_bits = W.
 * Code lengths above 17 (the encoder's maximum, huffman.cu:321), symbols of
 * 258 or more and selectors of n_groups or more give BZG_ERR_INVALID. */
bzg_status bzg_huff_encode(const uint16_t *d_sym, int32_t s,
                           const uint8_t *h_len, const int32_t *h_code,
                           const uint8_t *d_sel, int32_t n_sel,
                           int32_t n_groups, uint32_t *d_words,
                           int64_t *h_bits, void *stream);

#ifdef __cplusplus
}
#endif
#endif
