"""ctypes binding of libbzg.so (C ABI: include/bzg.h).

Stage functions take and return torch tensors on the current HIP device,
mirroring the reference's stage API (include/compression.h of
SebastianBoteroLeonik/bzip2gpu); whole-path functions take and return bytes.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libbzg.so")
REF_BLOCK_BYTES = 900000

_lib = None


class BzgError(RuntimeError):
    pass


class BzgDataError(BzgError):
    """Malformed or corrupt .bz2 input (status BZG_ERR_DATA)."""


ERR_CAPACITY = 3
ERR_DATA = 5


def build(jobs: int = 8) -> None:
    """Compiles libbzg.so and the bzip2gpu CLI for gfx950 (in-tree)."""
    subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"], check=True)


def library_path() -> str:
    return _SO


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_SO):
        raise BzgError(f"{_SO} is not built; run bzip2gpu_amd.build() (hipcc, gfx950)")
    L = ctypes.CDLL(_SO)
    P, i32, i64, u32, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_size_t
    I32P, I64P, U32P, SZP = (ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                             ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_size_t))
    sig = {
        "bzg_version": ([], ctypes.c_char_p),
        "bzg_last_error": ([], ctypes.c_char_p),
        "bzg_init": ([ctypes.c_int], i32),
        "bzg_compress_bound": ([sz, i32], sz),
        "bzg_compress_buffer": ([P, sz, ctypes.c_int, i32, P, sz, SZP], i32),
        "bzg_compress_plain": ([P, sz, ctypes.c_int, i32, P, sz, SZP], i32),
        "bzg_compress_plain_device": ([P, sz, ctypes.c_int, i32, P, sz, SZP, P], i32),
        "bzg_compress_batch_device": ([P, i32, P, P, P, i32, ctypes.POINTER(ctypes.c_void_p), P, P, P], i32),
        "bzg_last_round_stats": ([P, P, i32], i32),
        "bzg_server_create": ([i32, i32, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)], i32),
        "bzg_server_destroy": ([P], None),
        "bzg_server_submit": ([P, P, ctypes.c_uint64, i32, I64P], i32),
        "bzg_server_wait": ([P, i64, I64P, I64P, I64P], i32),
        "bzg_server_keep": ([P, P, ctypes.c_uint64], i32),
        "bzg_server_replay": ([P, P, i32, P, P, P, P, P, P, P, P], i32),
        "bzg_set_max_batch_bytes": ([ctypes.c_uint64], i32),
        "bzg_set_batch_overlap": ([i32], i32),
        "bzg_last_stage_times": ([P, I32P, I64P], i32),
        "bzg_last_sort_rows": ([I64P, I64P, I64P], i32),
        "bzg_tune_force": ([i32, i32], i32),
        "bzg_set_stage_graphs": ([i32], i32),
        "bzg_graph_captures": ([], i64),
        "bzg_graph_replays": ([], i64),
        "bzg_tune_state": ([i32, i32, P, P, P, i32], i32),
        "bzg_probe_kernel": ([ctypes.c_char_p], i32),
        "bzg_probe_read": ([ctypes.POINTER(ctypes.c_float), I32P], i32),
        "bzg_crc32": ([P, i32, U32P, P], i32),
        "bzg_rle1_bound": ([i32], i32),
        "bzg_rle1": ([P, i32, P, I32P, P], i32),
        "bzg_bwt": ([P, i32, P, I32P, P], i32),
        "bzg_symbols_table": ([P, i32, P, I32P, P], i32),
        "bzg_mtf": ([P, P, i32, P, I32P, P, I32P, P], i32),
        "bzg_rle2": ([P, i32, i32, P, P, P], i32),
        "bzg_histogram": ([P, i32, P, P], i32),
        "bzg_huff_build": ([P, i32, i32, P, P, P, I32P, I32P, P], i32),
        "bzg_huff_encode": ([P, i32, P, P, P, i32, i32, P, I64P, P], i32),
        "bzg_decompress_buffer": ([P, sz, P, sz, SZP], i32),
        "bzg_decompress_fetch": ([P, sz, SZP], i32),
        "bzg_decompress_device": ([P, sz, ctypes.POINTER(ctypes.c_void_p), SZP, P], i32),
        "bzg_last_decode_times": ([P, I32P, I32P], i32),
        "bzg_compress_batch_device_noalloc": ([P, i32, P, P, P, i32, P, sz, P, P, P], i32),
        "bzg_compress_segment": ([P, sz, i32, i32, U32P, P, sz, SZP, I64P, P], i32),
        "bzg_compress_plain_threads": ([P, i32, P, P, P, i32, i32, P, sz, P, P], i32),
        "bzg_last_batch_totals": ([P], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(st: int, what: str) -> None:
    if st != 0:
        msg = lib().bzg_last_error().decode(errors="replace")
        cls = BzgDataError if st == ERR_DATA else BzgError
        raise cls(f"{what} failed (status {st}): {msg}")


def version() -> str:
    return lib().bzg_version().decode()


def _np_u8(data) -> np.ndarray:
    if isinstance(data, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(data), dtype=np.uint8)
    return np.ascontiguousarray(data, dtype=np.uint8)


def compress_bound(n: int, block_bytes: int = REF_BLOCK_BYTES) -> int:
    return int(lib().bzg_compress_bound(n, block_bytes))


def _compress(fn, data, level, block_bytes) -> bytes:
    d = _np_u8(data)
    cap = compress_bound(d.size, block_bytes)
    out = np.empty(cap, dtype=np.uint8)
    n = ctypes.c_size_t(0)
    st = fn(d.ctypes.data_as(ctypes.c_void_p) if d.size else None, d.size, level, block_bytes,
            out.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(n))
    _check(st, fn.__name__)
    return out[: n.value].tobytes()


def compress(data, level: int = 9, block_bytes: int = REF_BLOCK_BYTES) -> bytes:
    """Batched device pipeline (all blocks at once, device-side assembly)."""
    return _compress(lib().bzg_compress_buffer, data, level, block_bytes)


def compress_plain(data, level: int = 9, block_bytes: int = REF_BLOCK_BYTES) -> bytes:
    """Plain path: one block per pipeline run, host-side serialisation."""
    return _compress(lib().bzg_compress_plain, data, level, block_bytes)


def compress_plain_device(d_in, level: int = 9, block_bytes: int = REF_BLOCK_BYTES, stream=None) -> bytes:
    """Plain path on an input already in device memory (torch uint8 tensor)."""
    n = d_in.numel()
    cap = compress_bound(n, block_bytes)
    out = np.empty(cap, dtype=np.uint8)
    got = ctypes.c_size_t(0)
    st = lib().bzg_compress_plain_device(ctypes.c_void_p(d_in.data_ptr()) if n else None, n, level, block_bytes,
                                         out.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(got),
                                         ctypes.c_void_p(stream) if stream else None)
    _check(st, "bzg_compress_plain_device")
    return out[: got.value].tobytes()


def compress_batch_device(d_in, offsets, lengths, levels, block_bytes: int = 0, stream=None):
    """Many independent inputs already on the device (torch uint8 tensor).

    Returns (device_out_ptr, byte_offsets np.int64[n], sizes np.int64[n]).
    The output buffer belongs to the library and is valid until the next call.
    """
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    ln = np.ascontiguousarray(lengths, dtype=np.uint64)
    lv = np.ascontiguousarray(levels, dtype=np.int32)
    n = off.size
    bo = np.zeros(n, dtype=np.int64)
    by = np.zeros(n, dtype=np.int64)
    dout = ctypes.c_void_p(0)
    st = lib().bzg_compress_batch_device(
        ctypes.c_void_p(d_in.data_ptr()), n, off.ctypes.data_as(ctypes.c_void_p),
        ln.ctypes.data_as(ctypes.c_void_p), lv.ctypes.data_as(ctypes.c_void_p), block_bytes,
        ctypes.byref(dout), bo.ctypes.data_as(ctypes.c_void_p), by.ctypes.data_as(ctypes.c_void_p),
        ctypes.c_void_p(stream) if stream else None)
    _check(st, "bzg_compress_batch_device")
    return dout.value, bo, by


def compress_plain_threads(d_in, offsets, lengths, levels, block_bytes: int = 0, threads: int = 16):
    """The reference's concurrency on the plain path: `threads` host threads,
    one pipeline and HIP stream each, one block per pipeline run; host
    serialisation.  Returns the list of streams (bytes)."""
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    ln = np.ascontiguousarray(lengths, dtype=np.uint64)
    lv = np.ascontiguousarray(levels, dtype=np.int32)
    n = off.size
    cap = sum(compress_bound(int(x), block_bytes if block_bytes > 0 else 100000 * int(v)) for x, v in zip(ln, lv))
    out = np.empty(max(1, cap), dtype=np.uint8)
    bo = np.zeros(n, dtype=np.int64)
    by = np.zeros(n, dtype=np.int64)
    c = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    _check(lib().bzg_compress_plain_threads(ctypes.c_void_p(d_in.data_ptr()), n, c(off), c(ln), c(lv), block_bytes,
                                            threads, c(out), out.size, c(bo), c(by)), "bzg_compress_plain_threads")
    return [out[int(bo[i]):int(bo[i] + by[i])].tobytes() for i in range(n)]


def last_batch_totals():
    """Totals of the last batched call: dict of blocks, n, m, s, bits, m_rounds."""
    t = np.zeros(6, np.int64)
    _check(lib().bzg_last_batch_totals(t.ctypes.data_as(ctypes.c_void_p)), "bzg_last_batch_totals")
    return dict(zip(["blocks", "n", "m", "s", "bits", "m_rounds"], [int(x) for x in t]))


def compress_batch_device_noalloc(d_in, offsets, lengths, levels, d_out, block_bytes: int = 0, stream=None):
    """Allocation-free batched path: the streams go to the caller's device
    tensor d_out (uint8); the library allocates nothing.  Returns
    (byte_offsets np.int64[n], sizes np.int64[n]) into d_out."""
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    ln = np.ascontiguousarray(lengths, dtype=np.uint64)
    lv = np.ascontiguousarray(levels, dtype=np.int32)
    n = off.size
    bo = np.zeros(n, dtype=np.int64)
    by = np.zeros(n, dtype=np.int64)
    st = lib().bzg_compress_batch_device_noalloc(
        ctypes.c_void_p(d_in.data_ptr()), n, off.ctypes.data_as(ctypes.c_void_p),
        ln.ctypes.data_as(ctypes.c_void_p), lv.ctypes.data_as(ctypes.c_void_p), block_bytes,
        ctypes.c_void_p(d_out.data_ptr()), d_out.numel(), bo.ctypes.data_as(ctypes.c_void_p),
        by.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(stream) if stream else None)
    _check(st, "bzg_compress_batch_device_noalloc")
    return bo, by


def decompress(data) -> bytes:
    """Decodes one or more concatenated .bz2 streams on the device (host
    bytes in, host bytes out).  Raises BzgDataError on malformed or corrupt
    input (bad header or magic, CRC mismatch, truncation)."""
    d = _np_u8(data)
    cap = max(4 * d.size, 1 << 20)
    out = np.empty(cap, dtype=np.uint8)
    n = ctypes.c_size_t(0)
    L = lib()
    st = L.bzg_decompress_buffer(d.ctypes.data_as(ctypes.c_void_p) if d.size else None, d.size,
                                 out.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(n))
    if st == ERR_CAPACITY:
        out = np.empty(max(n.value, 1), dtype=np.uint8)
        st = L.bzg_decompress_fetch(out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(n))
    _check(st, "bzg_decompress_buffer")
    return out[: n.value].tobytes()


def decompress_device(d_in, stream=None):
    """Decodes .bz2 data held in a device uint8 tensor.  Returns (device
    address, length) of the raw output, a library-owned buffer valid until
    the next decompression call."""
    ptr = ctypes.c_void_p(0)
    n = ctypes.c_size_t(0)
    st = lib().bzg_decompress_device(_ptr(d_in) if d_in.numel() else None, d_in.numel(), ctypes.byref(ptr),
                                     ctypes.byref(n), stream if stream is not None else _stream())
    _check(st, "bzg_decompress_device")
    return (ptr.value or 0), n.value


def last_decode_times():
    """(times_ms[5], blocks, streams) of the last decompression: Huffman
    (incl. magic scan and chain walk), inverse MTF/RLE2, FL-map sort,
    inverse BWT walks, inverse RLE1 + CRC."""
    ms = (ctypes.c_float * 5)()
    nb = ctypes.c_int32(0)
    ns = ctypes.c_int32(0)
    _check(lib().bzg_last_decode_times(ms, ctypes.byref(nb), ctypes.byref(ns)), "bzg_last_decode_times")
    return list(ms), nb.value, ns.value


def last_round_stats(max_rounds: int = 64):
    """[(rows, gpu ms)] of every suffix-sort round of the last pipeline run."""
    rows = np.zeros(max_rounds, dtype=np.int64)
    ms = np.zeros(max_rounds, dtype=np.float32)
    n = lib().bzg_last_round_stats(rows.ctypes.data_as(ctypes.c_void_p), ms.ctypes.data_as(ctypes.c_void_p), max_rounds)
    return [(int(rows[i]), float(ms[i])) for i in range(max(0, min(n, max_rounds)))]


def set_max_batch_bytes(n: int) -> None:
    _check(lib().bzg_set_max_batch_bytes(int(n)), "bzg_set_max_batch_bytes")


def set_batch_overlap(on: bool) -> None:
    """Overlap consecutive sub-batches on two pipelines; see bzg_set_batch_overlap."""
    _check(lib().bzg_set_batch_overlap(1 if on else 0), "bzg_set_batch_overlap")


def last_stage_times():
    """({stage: ms}, bwt rounds, suffix-sort rows processed) of the last batched call."""
    ms = (ctypes.c_float * 6)()
    r = ctypes.c_int32(0)
    rows = ctypes.c_int64(0)
    _check(lib().bzg_last_stage_times(ms, ctypes.byref(r), ctypes.byref(rows)), "bzg_last_stage_times")
    keys = ["front", "bwt", "mtf_rle2", "huffman", "pack", "assemble"]
    return dict(zip(keys, list(ms))), int(r.value), int(rows.value)


def last_sort_rows():
    """(small-group rows, large-group rows, large-group rows x radix passes)
    of the last batched call, summed over suffix-sort rounds."""
    a, b, c = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    _check(lib().bzg_last_sort_rows(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "bzg_last_sort_rows")
    return int(a.value), int(b.value), int(c.value)


def set_stage_graphs(on: bool) -> None:
    """Run the non-sort stage groups as replayed HIP graphs (default) or eagerly."""
    _check(lib().bzg_set_stage_graphs(1 if on else 0), "bzg_set_stage_graphs")


def graph_counts():
    """(stage-group captures, graph launches) of the device context."""
    return int(lib().bzg_graph_captures()), int(lib().bzg_graph_replays())


def tune_force(knob: int, variant: int) -> None:
    """Pin an autotuned launch shape (-1: automatic); see bzg_tune_force."""
    _check(lib().bzg_tune_force(int(knob), int(variant)), "bzg_tune_force")


def tune_state(knob: int = 0, size_class: int = 0):
    """(chosen variant or -1, [(launch value, mean ns per unit, samples)])."""
    vals = np.full(4, -1, dtype=np.int32)
    ns = np.zeros(4, dtype=np.float32)
    n = np.zeros(4, dtype=np.int32)
    c = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    chosen = lib().bzg_tune_state(int(knob), int(size_class), c(vals), c(ns), c(n), 4)
    if chosen == -2:
        raise BzgError("bzg_tune_state: bad argument")
    return int(chosen), [(int(vals[i]), float(ns[i]), int(n[i])) for i in range(4) if vals[i] >= 0]


def probe_kernel(name) -> None:
    """Time every later launch of kernel `name` with HIP events (None: off)."""
    _check(lib().bzg_probe_kernel(name.encode() if name else None), "bzg_probe_kernel")


def probe_read():
    """(total ms, launches) of the probed kernel since the last read."""
    ms = ctypes.c_float(0)
    n = ctypes.c_int32(0)
    _check(lib().bzg_probe_read(ctypes.byref(ms), ctypes.byref(n)), "bzg_probe_read")
    return float(ms.value), int(n.value)


# ------------------------------------------------------------- stage level
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise BzgError("no HIP device available")
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def crc32(d_in) -> int:
    c = ctypes.c_uint32(0)
    _check(lib().bzg_crc32(_ptr(d_in), d_in.numel(), ctypes.byref(c), _stream()), "bzg_crc32")
    return int(c.value)


def rle1(d_in):
    torch = _torch()
    n = d_in.numel()
    out = torch.empty(max(1, lib().bzg_rle1_bound(n)), dtype=torch.uint8, device=d_in.device)
    m = ctypes.c_int32(0)
    _check(lib().bzg_rle1(_ptr(d_in), n, _ptr(out), ctypes.byref(m), _stream()), "bzg_rle1")
    return out[: m.value]


def bwt(d_in):
    torch = _torch()
    m = d_in.numel()
    sa = torch.empty(max(1, m), dtype=torch.int32, device=d_in.device)
    r = ctypes.c_int32(0)
    _check(lib().bzg_bwt(_ptr(d_in), m, _ptr(sa), ctypes.byref(r), _stream()), "bzg_bwt")
    return sa[:m], int(r.value)


def symbols_table(d_in):
    tab = np.zeros(256, dtype=np.uint8)
    c = ctypes.c_int32(0)
    _check(lib().bzg_symbols_table(_ptr(d_in), d_in.numel(), tab.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.byref(c), _stream()), "bzg_symbols_table")
    return tab[: c.value].copy()


def mtf(d_rle1, d_sa):
    torch = _torch()
    m = d_rle1.numel()
    out = torch.empty(max(1, m), dtype=torch.uint8, device=d_rle1.device)
    op = ctypes.c_int32(0)
    nu = ctypes.c_int32(0)
    pres = np.zeros(256, dtype=np.uint8)
    _check(lib().bzg_mtf(_ptr(d_rle1), _ptr(d_sa), m, _ptr(out), ctypes.byref(op),
                         pres.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nu), _stream()), "bzg_mtf")
    return out[:m], int(op.value), int(nu.value), pres.astype(bool)


def rle2(d_mtf, table_size: int):
    torch = _torch()
    m = d_mtf.numel()
    out = torch.empty(m + 2, dtype=torch.int16, device=d_mtf.device)
    n = torch.zeros(1, dtype=torch.int32, device=d_mtf.device)
    _check(lib().bzg_rle2(_ptr(d_mtf), m, table_size, _ptr(out), _ptr(n), _stream()), "bzg_rle2")
    s = int(n.item())
    return out[:s]


def histogram(d_sym, d_freq=None):
    torch = _torch()
    if d_freq is None:
        d_freq = torch.zeros(258, dtype=torch.int32, device=d_sym.device)
    _check(lib().bzg_histogram(_ptr(d_sym), d_sym.numel(), _ptr(d_freq), _stream()), "bzg_histogram")
    return d_freq


def huff_build(d_sym, alphabet_size: int):
    s = d_sym.numel()
    ln = np.zeros((6, 258), dtype=np.uint8)
    cd = np.zeros((6, 258), dtype=np.int32)
    sel = np.zeros(s // 50 + 2, dtype=np.uint8)
    ng = ctypes.c_int32(0)
    ns = ctypes.c_int32(0)
    _check(lib().bzg_huff_build(_ptr(d_sym), s, alphabet_size, ln.ctypes.data_as(ctypes.c_void_p),
                                cd.ctypes.data_as(ctypes.c_void_p), sel.ctypes.data_as(ctypes.c_void_p),
                                ctypes.byref(ng), ctypes.byref(ns), _stream()), "bzg_huff_build")
    return ln, cd, sel[: ns.value].copy(), int(ng.value)


def huff_encode(d_sym, ln, cd, sel, n_groups: int):
    torch = _torch()
    s = d_sym.numel()
    ln = np.ascontiguousarray(ln, dtype=np.uint8)
    cd = np.ascontiguousarray(cd, dtype=np.int32)
    d_sel = torch.from_numpy(np.ascontiguousarray(sel, dtype=np.uint8)).to(d_sym.device)
    words = torch.zeros(s * 17 // 32 + 4, dtype=torch.int32, device=d_sym.device)
    bits = ctypes.c_int64(0)
    _check(lib().bzg_huff_encode(_ptr(d_sym), s, ln.ctypes.data_as(ctypes.c_void_p),
                                 cd.ctypes.data_as(ctypes.c_void_p), _ptr(d_sel), int(d_sel.numel()),
                                 n_groups, _ptr(words), ctypes.byref(bits), _stream()), "bzg_huff_encode")
    nw = (bits.value + 31) // 32
    return words[:nw], int(bits.value)


class Server:
    """Continuous-batching request loop (C++), see bzg_server_* in bzg.h.

    mode "batched": queued requests that fit are compressed together, one
    pipeline run per worker at a time; mode "plain": each request alone on
    the per-block path.
    """

    def __init__(self, mode: str = "batched", workers: int = 4, max_batch_bytes: int = 256 << 20,
                 hbm_budget_bytes: int = 200 * 10**9):
        h = ctypes.c_void_p(0)
        _check(lib().bzg_server_create(0 if mode == "batched" else 1, workers, max_batch_bytes, hbm_budget_bytes,
                                       ctypes.byref(h)), "bzg_server_create")
        self._h = h

    def close(self):
        if self._h:
            lib().bzg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def keep_outputs(self, d_arena):
        """Keep every output in a torch uint8 device tensor (None: sizes only)."""
        _check(lib().bzg_server_keep(self._h, _ptr(d_arena) if d_arena is not None else None,
                                     d_arena.numel() if d_arena is not None else 0), "bzg_server_keep")

    def submit(self, d_buf, level: int) -> int:
        t = ctypes.c_int64(0)
        _check(lib().bzg_server_submit(self._h, _ptr(d_buf), d_buf.numel(), level, ctypes.byref(t)),
               "bzg_server_submit")
        return int(t.value)

    def wait(self, ticket: int):
        """(output bytes, arena offset or -1, latency ns)."""
        b, o, l = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        _check(lib().bzg_server_wait(self._h, ticket, ctypes.byref(b), ctypes.byref(o), ctypes.byref(l)),
               "bzg_server_wait")
        return int(b.value), int(o.value), int(l.value)

    def replay(self, d_pool, offsets, lengths, levels, arrival_ns):
        """Open-loop replay; returns (done_ns, out_bytes, out_off, batch_size) arrays."""
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        ln = np.ascontiguousarray(lengths, dtype=np.uint64)
        lv = np.ascontiguousarray(levels, dtype=np.int32)
        arr = np.ascontiguousarray(arrival_ns, dtype=np.int64)
        n = off.size
        done = np.zeros(n, np.int64)
        ob = np.zeros(n, np.int64)
        oo = np.zeros(n, np.int64)
        bs = np.zeros(n, np.int32)
        c = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _check(lib().bzg_server_replay(self._h, ctypes.c_void_p(d_pool.data_ptr()), n, c(off), c(ln), c(lv), c(arr),
                                       c(done), c(ob), c(oo), c(bs)), "bzg_server_replay")
        return done, ob, oo, bs
