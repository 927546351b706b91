"""Independent pure-Python statements of the behaviour the reference's own
tests assert (the CPU functions inside tests/compression/*.cu of
SebastianBoteroLeonik/bzip2gpu).  They pin the C oracle; they are written
from the stated rules, loop by loop, for small and medium inputs.
"""
import numpy as np


def rle1(data: bytes) -> bytes:
    """tests/compression/rle1.cu:8-32: runs capped at 259; a run r < 4 is
    copied, else 4 copies + (r - 4)."""
    out = bytearray()
    n = len(data)
    run = 1
    for i in range(n):
        if i + 1 < n and data[i] == data[i + 1] and run < 259:
            run += 1
            continue
        if run < 4:
            out += bytes([data[i]]) * run
        else:
            out += bytes([data[i]]) * 4 + bytes([run - 4])
        run = 1
    return bytes(out)


def bwt_bruteforce(data: bytes) -> np.ndarray:
    """tests/compression/fbwt.cu:12-24: sort rotation start indices by the
    rotation's bytes, ties by smaller index."""
    n = len(data)
    dd = data + data
    idx = sorted(range(n), key=lambda a: (dd[a:a + n], a))
    return np.array(idx, dtype=np.int32)


def mtf(data: bytes) -> np.ndarray:
    """tests/compression/fmtf.cu:48-74: list starts as the sorted set of
    present bytes; output = position, then move to front."""
    lst = sorted(set(data))
    out = np.empty(len(data), dtype=np.uint8)
    for i, c in enumerate(data):
        p = lst.index(c)
        out[i] = p
        if p:
            del lst[p]
            lst.insert(0, c)
    return out


def rle2(data: bytes, table_size: int) -> np.ndarray:
    """tests/compression/rle2.cu:9-33: nonzero v -> v+1; a zero run of k ->
    k--, emit k&1, k >>= 1 until k == 0; EOB = table_size + 1."""
    out = []
    i, n = 0, len(data)
    while i < n:
        if data[i]:
            out.append(data[i] + 1)
            i += 1
        else:
            k = 0
            while i < n and data[i] == 0:
                k += 1
                i += 1
            while k > 0:
                k -= 1
                out.append(k & 1)
                k >>= 1
    out.append(table_size + 1)
    return np.array(out, dtype=np.uint16)


def inverse_bwt(last: bytes, orig_ptr: int) -> bytes:
    """Textbook inverse BWT (LF mapping) of cyclic rotations."""
    n = len(last)
    order = sorted(range(n), key=lambda i: (last[i], i))
    t = [0] * n
    for j, i in enumerate(order):
        t[j] = i
    out = bytearray(n)
    p = t[orig_ptr]
    for k in range(n):
        out[k] = last[p]
        p = t[p]
    return bytes(out)


def decode_huffman(words: np.ndarray, bits: int, lens, codes, sel, alpha: int) -> np.ndarray:
    """Decode MSB-first packed symbols with the table of each 50-symbol group
    (tests/compression/huffman.cu:86-159)."""
    tabs = []
    for t in range(lens.shape[0]):
        tabs.append({(int(lens[t][v]), int(codes[t][v])): v for v in range(alpha)})
    out = []
    pos = 0
    stream = np.unpackbits(np.asarray(words, dtype=">u4").view(np.uint8))
    while pos < bits:
        t = int(sel[len(out) // 50])
        code, ln = 0, 0
        while True:
            code = (code << 1) | int(stream[pos])
            pos += 1
            ln += 1
            if (ln, code) in tabs[t]:
                out.append(tabs[t][(ln, code)])
                break
            if ln > 20:
                raise ValueError("undecodable")
    return np.array(out, dtype=np.uint16)
