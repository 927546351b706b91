"""ctypes bindings to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
(bzip2gpu_amd/libbzg.so) never loads this.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

MAX_GROUPS = 6
MAX_ALPHA = 258


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Block(ctypes.Structure):
    _fields_ = [
        ("crc", ctypes.c_uint32),
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("s", ctypes.c_int32),
        ("orig_ptr", ctypes.c_int32),
        ("n_in_use", ctypes.c_int32),
        ("alpha_size", ctypes.c_int32),
        ("present", ctypes.c_uint8 * 256),
        ("n_groups", ctypes.c_int32),
        ("n_sel", ctypes.c_int32),
        ("len", (ctypes.c_uint8 * MAX_ALPHA) * MAX_GROUPS),
        ("code", (ctypes.c_int32 * MAX_ALPHA) * MAX_GROUPS),
        ("bits", ctypes.c_int64),
        ("rounds", ctypes.c_int32),
        ("sel", ctypes.POINTER(ctypes.c_uint8)),
        ("sel_mtf", ctypes.POINTER(ctypes.c_uint8)),
        ("words", ctypes.POINTER(ctypes.c_uint32)),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        i32, i64, u32, sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_size_t
        L.orc_crc32.argtypes = [P, sz]; L.orc_crc32.restype = u32
        L.orc_crc32_update.argtypes = [u32, P, sz]; L.orc_crc32_update.restype = u32
        L.orc_rle1.argtypes = [P, i32, P]; L.orc_rle1.restype = i32
        L.orc_bwt.argtypes = [P, i32, P]; L.orc_bwt.restype = i32
        L.orc_mtf.argtypes = [P, P, i32, P, P, P]; L.orc_mtf.restype = i32
        L.orc_mtf_plain.argtypes = [P, i32, P]; L.orc_mtf_plain.restype = None
        L.orc_rle2.argtypes = [P, i32, i32, P]; L.orc_rle2.restype = i32
        L.orc_huff_build.argtypes = [P, i32, i32, P, P, P, P]; L.orc_huff_build.restype = i32
        L.orc_make_code_lengths.argtypes = [P, P, i32, i32]; L.orc_make_code_lengths.restype = None
        L.orc_selector_mtf.argtypes = [P, i32, i32, P]; L.orc_selector_mtf.restype = None
        L.orc_huff_encode.argtypes = [P, i32, P, P, P, P]; L.orc_huff_encode.restype = i64
        L.orc_compress_block.argtypes = [P, i32, ctypes.POINTER(Block)]; L.orc_compress_block.restype = ctypes.c_int
        L.orc_block_free.argtypes = [ctypes.POINTER(Block)]; L.orc_block_free.restype = None
        L.orc_compress.argtypes = [P, sz, ctypes.c_int, i32, P, sz]; L.orc_compress.restype = i64
        L.orc_compress_mt.argtypes = [P, sz, ctypes.c_int, i32, P, sz, ctypes.c_int]; L.orc_compress_mt.restype = i64
        L.orc_compress_bound.argtypes = [sz]; L.orc_compress_bound.restype = sz
        _lib = L
    return _lib


def _u8(a) -> np.ndarray:
    if isinstance(a, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(a), dtype=np.uint8)
    return np.ascontiguousarray(a, dtype=np.uint8)


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def crc32(data) -> int:
    d = _u8(data)
    return int(lib().orc_crc32(_p(d), d.size))


def rle1(data) -> np.ndarray:
    d = _u8(data)
    out = np.zeros(5 * ((d.size + 3) // 4) + 8, dtype=np.uint8)
    m = lib().orc_rle1(_p(d), d.size, _p(out))
    return out[:m].copy()


def bwt(data):
    """Returns (suffix array int32[m], rounds)."""
    d = _u8(data)
    sa = np.zeros(max(d.size, 1), dtype=np.int32)
    r = lib().orc_bwt(_p(d), d.size, _p(sa))
    return sa[: d.size].copy(), int(r)


def mtf(rle1_bytes, sa):
    """Returns (mtf u8[m], orig_ptr, n_in_use, present bool[256])."""
    d = _u8(rle1_bytes)
    s = np.ascontiguousarray(sa, dtype=np.int32)
    out = np.zeros(max(d.size, 1), dtype=np.uint8)
    op = ctypes.c_int32(0)
    pres = np.zeros(256, dtype=np.uint8)
    n = lib().orc_mtf(_p(d), _p(s), d.size, _p(out), ctypes.byref(op), _p(pres))
    return out[: d.size].copy(), int(op.value), int(n), pres.astype(bool)


def mtf_plain(data) -> np.ndarray:
    d = _u8(data)
    out = np.zeros(max(d.size, 1), dtype=np.uint8)
    lib().orc_mtf_plain(_p(d), d.size, _p(out))
    return out[: d.size].copy()


def rle2(mtf_bytes, n_in_use: int) -> np.ndarray:
    d = _u8(mtf_bytes)
    out = np.zeros(d.size + 2, dtype=np.uint16)
    s = lib().orc_rle2(_p(d), d.size, n_in_use, _p(out))
    return out[:s].copy()


def huff_build(sym, alphabet_size: int):
    """Returns (len u8[6,258], code i32[6,258], sel u8[n_sel], n_groups)."""
    d = np.ascontiguousarray(sym, dtype=np.uint16)
    ln = np.zeros((MAX_GROUPS, MAX_ALPHA), dtype=np.uint8)
    cd = np.zeros((MAX_GROUPS, MAX_ALPHA), dtype=np.int32)
    sel = np.zeros(d.size // 50 + 2, dtype=np.uint8)
    ng = ctypes.c_int32(0)
    ns = lib().orc_huff_build(_p(d), d.size, alphabet_size, _p(ln), _p(cd), _p(sel), ctypes.byref(ng))
    return ln, cd, sel[:ns].copy(), int(ng.value)


def make_code_lengths(freq, alpha: int, max_len: int = 17) -> np.ndarray:
    f = np.ascontiguousarray(freq, dtype=np.int32)
    out = np.zeros(MAX_ALPHA, dtype=np.uint8)
    lib().orc_make_code_lengths(_p(f), _p(out), alpha, max_len)
    return out[:alpha].copy()


def selector_mtf(sel, n_groups: int) -> np.ndarray:
    s = _u8(sel)
    out = np.zeros(max(s.size, 1), dtype=np.uint8)
    lib().orc_selector_mtf(_p(s), s.size, n_groups, _p(out))
    return out[: s.size].copy()


def huff_encode(sym, ln, cd, sel):
    d = np.ascontiguousarray(sym, dtype=np.uint16)
    ln = np.ascontiguousarray(ln, dtype=np.uint8)
    cd = np.ascontiguousarray(cd, dtype=np.int32)
    s = _u8(sel)
    words = np.zeros(d.size * 17 // 32 + 4, dtype=np.uint32)
    bits = lib().orc_huff_encode(_p(d), d.size, _p(ln), _p(cd), _p(s), _p(words))
    return words[: (bits + 31) // 32].copy(), int(bits)


def compress_block(data) -> dict:
    d = _u8(data)
    b = Block()
    rc = lib().orc_compress_block(_p(d), d.size, ctypes.byref(b))
    if rc != 0:
        raise ValueError("empty block")
    nw = (b.bits + 31) // 32
    res = {
        "crc": b.crc, "n": b.n, "m": b.m, "s": b.s, "orig_ptr": b.orig_ptr,
        "n_in_use": b.n_in_use, "alpha_size": b.alpha_size,
        "present": np.frombuffer(bytes(b.present), dtype=np.uint8).astype(bool),
        "n_groups": b.n_groups, "n_sel": b.n_sel,
        "len": np.array([[b.len[g][v] for v in range(MAX_ALPHA)] for g in range(MAX_GROUPS)], dtype=np.uint8),
        "code": np.array([[b.code[g][v] for v in range(MAX_ALPHA)] for g in range(MAX_GROUPS)], dtype=np.int32),
        "bits": b.bits, "rounds": b.rounds,
        "sel": np.ctypeslib.as_array(b.sel, shape=(b.n_sel,)).copy() if b.n_sel else np.zeros(0, np.uint8),
        "sel_mtf": np.ctypeslib.as_array(b.sel_mtf, shape=(b.n_sel,)).copy() if b.n_sel else np.zeros(0, np.uint8),
        "words": np.ctypeslib.as_array(b.words, shape=(nw,)).copy() if nw else np.zeros(0, np.uint32),
    }
    lib().orc_block_free(ctypes.byref(b))
    return res


def compress(data, level: int = 9, block_bytes: int = 900000, threads: int = 1) -> bytes:
    d = _u8(data)
    cap = int(lib().orc_compress_bound(d.size))
    out = np.zeros(cap, dtype=np.uint8)
    n = lib().orc_compress_mt(_p(d), d.size, level, block_bytes, _p(out), cap, threads)
    if n < 0:
        raise RuntimeError("oracle output overflow")
    return out[:n].tobytes()
