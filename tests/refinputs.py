"""The reference's seeded test inputs, regenerated bit-for-bit.

The reference's tests (tests/compression/*.cu of SebastianBoteroLeonik/
bzip2gpu) draw every input from glibc srand(2137)/rand(); glibc's rand() is
called here through ctypes, so the byte sequences are the same ones its
tests assert on.  Each function cites the test macro it regenerates.
"""
import ctypes
import functools

import numpy as np

_libc = ctypes.CDLL("libc.so.6")
_libc.rand.restype = ctypes.c_int
_libc.srand.argtypes = [ctypes.c_uint]


def _rands(seed, count):
    _libc.srand(seed)
    r = _libc.rand
    return [r() for _ in range(count)]


@functools.lru_cache(maxsize=None)
def rle1_input(max_run: int) -> bytes:
    """RUN_RLE1(MAX_RUN_LENGTH), tests/compression/rle1.cu:34-47."""
    _libc.srand(2137)
    r = _libc.rand
    out = bytearray(1000000)
    chr_ = 0
    reps = 0
    for i in range(1000000):
        if not reps:
            reps = r() % max_run + 1
            chr_ = (chr_ + r() % 100 + 1) & 0xFF
        out[i] = chr_
        reps -= 1
    return bytes(out)


@functools.lru_cache(maxsize=None)
def bwt_input(alphabet: int) -> bytes:
    """RUN_FBWT(ALPHABET_SIZE), tests/compression/fbwt.cu:26-31."""
    return bytes(v % alphabet for v in _rands(2137, 5000))


@functools.lru_cache(maxsize=None)
def mtf_input(alphabet: int) -> bytes:
    """RUN_FMTF(ALPHABET_SIZE), tests/compression/fmtf.cu:75-80 (SA = (i+1) mod N)."""
    return bytes(v % alphabet for v in _rands(2137, 1000000))


@functools.lru_cache(maxsize=None)
def symbols_table_case():
    """mtf_symbol_table_generation, tests/compression/fmtf.cu:21-33."""
    _libc.srand(2137)
    r = _libc.rand
    table = [i for i in range(256) if r() % 3 == 0]
    data = bytes(table[r() % len(table)] for _ in range(100000))
    return table, data


@functools.lru_cache(maxsize=None)
def rle2_input(max_run: int) -> bytes:
    """RUN_RLE2(MAX_RUN_LENGTH), tests/compression/rle2.cu:35-48."""
    _libc.srand(2137)
    r = _libc.rand
    out = bytearray(1000000)
    reps = 0
    for i in range(1000000):
        if not reps:
            reps = r() % max_run + 1
            out[i] = r() % 256
        else:
            out[i] = out[i - 1]
        reps -= 1
    return bytes(out)


@functools.lru_cache(maxsize=None)
def rle2_incompressible_input() -> bytes:
    """rle2_incompressible, tests/compression/rle2.cu:89-95."""
    return bytes(v % 255 + 1 for v in _rands(2137, 1000000))


def table_size(data: bytes) -> int:
    return int(np.count_nonzero(np.bincount(np.frombuffer(data, np.uint8), minlength=256)))


@functools.lru_cache(maxsize=None)
def histogram_input() -> np.ndarray:
    """huffman_count_frequencies, tests/compression/huffman.cu:15-21."""
    return np.array([(v % 300) % 258 for v in _rands(2137, 10000)], dtype=np.uint16)


def stability_inputs():
    """huffman_tree_builder_stability, tests/compression/huffman.cu:49-55:
    one srand(2137), then 10000 symbols per alphabet size 1..255."""
    _libc.srand(2137)
    r = _libc.rand
    for a in range(1, 256):
        yield a, np.array([r() % a for _ in range(10000)], dtype=np.uint16)


@functools.lru_cache(maxsize=None)
def encoding_input() -> np.ndarray:
    """huffman_encoding, tests/compression/huffman.cu:171-179 (alphabet 15)."""
    _libc.srand(2137)
    r = _libc.rand
    out = np.zeros(10000, dtype=np.uint16)
    shift = 0
    for i in range(10000):
        if i % 50 == 0:
            shift = r() % 6
        out[i] = r() % (15 // 2) + shift
    return out
