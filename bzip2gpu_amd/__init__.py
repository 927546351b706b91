"""bzip2gpu_amd - bzip2 block compression on AMD MI355X (gfx950).

The compute path is libbzg.so: hand-written HIP kernels behind the C ABI in
include/bzg.h.  This package is a thin ctypes layer over it for Python
callers, tests and benchmarks.  There is no CPU fallback: if the library or
a GPU is missing, calls raise.
"""
from ._lib import (  # noqa: F401
    BzgDataError,
    BzgError,
    REF_BLOCK_BYTES,
    Server,
    build,
    bwt,
    compress,
    compress_batch_device,
    compress_batch_device_noalloc,
    compress_plain_threads,
    last_batch_totals,
    compress_bound,
    compress_plain,
    compress_plain_device,
    crc32,
    decompress,
    decompress_device,
    histogram,
    huff_build,
    huff_encode,
    lib,
    library_path,
    last_decode_times,
    last_round_stats,
    last_sort_rows,
    last_stage_times,
    mtf,
    probe_kernel,
    probe_read,
    rle1,
    rle2,
    set_max_batch_bytes,
    set_batch_overlap,
    symbols_table,
    set_stage_graphs,
    graph_counts,
    tune_force,
    tune_state,
    version,
)
