"""A small, independent bzip2 stream decoder for tests.

Stock libbz2 rejects any block whose RLE1 length exceeds 100000*level
bytes.  The reference cuts every block at 900000 raw bytes, and RLE1 can
grow a block (a run of exactly four bytes becomes five), so many valid
reference-format streams are refused by stock bzip2 for their size alone.
This decoder follows the published bzip2 format (Huffman tables with
50-symbol selectors, RUNA/RUNB zero runs, MTF, inverse BWT, RLE1 with a
count byte after four equal bytes) without that size limit, and checks
every block CRC and the stream CRC.
"""
import numpy as np

_MAGIC_BLOCK = 0x314159265359
_MAGIC_EOS = 0x177245385090

_CRC = []
for _i in range(256):
    _c = _i << 24
    for _ in range(8):
        _c = ((_c << 1) ^ 0x04C11DB7) & 0xFFFFFFFF if _c & 0x80000000 else (_c << 1) & 0xFFFFFFFF
    _CRC.append(_c)
_CRC = np.array(_CRC, dtype=np.uint32)


def crc32(data: bytes) -> int:
    c = 0xFFFFFFFF
    for b in data:
        c = ((c << 8) & 0xFFFFFFFF) ^ int(_CRC[(c >> 24) ^ b])
    return c ^ 0xFFFFFFFF


class _Bits:
    def __init__(self, data: bytes):
        self.bits = np.unpackbits(np.frombuffer(data, np.uint8))
        self.pos = 0
        b = np.concatenate([self.bits, np.zeros(40, np.uint8)]).astype(np.uint32)
        w = np.zeros(self.bits.size, dtype=np.uint32)
        for k in range(20):
            w = (w << 1) | b[k:k + self.bits.size]
        self.win20 = w  # 20-bit window starting at each bit position

    def get(self, n: int) -> int:
        v = 0
        for _ in range(n):
            v = (v << 1) | int(self.bits[self.pos])
            self.pos += 1
        return v


def _decode_block(br: _Bits, level: int):
    crc = br.get(32)
    if br.get(1):
        raise ValueError("randomised blocks not supported")
    orig_ptr = br.get(24)
    used = br.get(16)
    present = []
    for r in range(16):
        if (used >> (15 - r)) & 1:
            m = br.get(16)
            for q in range(16):
                if (m >> (15 - q)) & 1:
                    present.append(r * 16 + q)
    n_in_use = len(present)
    alpha = n_in_use + 2
    n_groups = br.get(3)
    n_sel = br.get(15)
    sel_mtf = []
    for _ in range(n_sel):
        j = 0
        while br.get(1):
            j += 1
        sel_mtf.append(j)
    lst = list(range(n_groups))
    sel = []
    for j in sel_mtf:
        v = lst.pop(j)
        lst.insert(0, v)
        sel.append(v)
    # coding tables -> 20-bit lookup (symbol, length) per table
    luts = []
    for _ in range(n_groups):
        cur = br.get(5)
        lens = []
        for _v in range(alpha):
            while br.get(1):
                cur += -1 if br.get(1) else 1
            lens.append(cur)
        lens = np.array(lens)
        code = 0
        lut_sym = np.zeros(1 << 20, dtype=np.int32)
        lut_len = np.zeros(1 << 20, dtype=np.int32)
        for ln in range(int(lens.min()), int(lens.max()) + 1):
            for v in range(alpha):
                if lens[v] == ln:
                    lo = code << (20 - ln)
                    hi = (code + 1) << (20 - ln)
                    lut_sym[lo:hi] = v
                    lut_len[lo:hi] = ln
                    code += 1
            code <<= 1
        luts.append((lut_sym, lut_len))
    # symbols
    syms = []
    eob = n_in_use + 1
    while True:
        g = len(syms) // 50
        ls, ll = luts[sel[g]]
        w = int(br.win20[br.pos])
        v = int(ls[w])
        br.pos += int(ll[w])
        syms.append(v)
        if v == eob:
            break
    # RLE2 + MTF inverse
    lst = list(present)
    out = bytearray()
    run, weight = 0, 1
    for v in syms[:-1]:
        if v <= 1:
            run += weight << v
            weight <<= 1
            continue
        if run:
            out += bytes([lst[0]]) * run
            run, weight = 0, 1
        c = lst.pop(v - 1)
        lst.insert(0, c)
        out.append(c)
    if run:
        out += bytes([lst[0]]) * run
    last = np.frombuffer(bytes(out), np.uint8)
    n = last.size
    if not (0 <= orig_ptr < n):
        raise ValueError("bad origPtr")
    t = np.argsort(last, kind="stable")
    res = np.empty(n, np.uint8)
    p = int(t[orig_ptr])
    for k in range(n):
        res[k] = last[p]
        p = int(t[p])
    # RLE1 inverse
    dec = bytearray()
    r = res.tobytes()
    i = 0
    while i < len(r):
        c = r[i]
        j = i
        while j < len(r) and j - i < 4 and r[j] == c:
            j += 1
        dec += r[i:j]
        if j - i == 4:
            dec += bytes([c]) * r[j]
            j += 1
        i = j
    if crc32(bytes(dec)) != crc:
        raise ValueError("block CRC mismatch")
    return bytes(dec), crc, n


def decompress(stream: bytes, max_block_rle1=None):
    """Decodes one bzip2 stream; returns (data, per-block RLE1 lengths)."""
    if stream[:3] != b"BZh" or not (0x31 <= stream[3] <= 0x39):
        raise ValueError("bad header")
    level = stream[3] - 0x30
    br = _Bits(stream)
    br.pos = 32
    out = bytearray()
    combined = 0
    sizes = []
    while True:
        magic = br.get(48)
        if magic == _MAGIC_EOS:
            break
        if magic != _MAGIC_BLOCK:
            raise ValueError("bad block magic")
        blk, crc, m = _decode_block(br, level)
        sizes.append(m)
        combined = ((combined << 1) | (combined >> 31)) & 0xFFFFFFFF
        combined ^= crc
        out += blk
    if br.get(32) != combined:
        raise ValueError("stream CRC mismatch")
    return bytes(out), sizes
