"""Pins the CPU oracle (the checker) against the reference's own test specs,
brute force, closed forms and stock libbz2.  CPU only."""
import bz2
import ctypes
import os

import numpy as np
import pytest

import oracle
from tests import refinputs, spec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ----------------------------------------------------------------- CRC
def test_crc_check_value():
    # CRC-32/BZIP2 check value of "123456789"
    assert oracle.crc32(b"123456789") == 0xFC891918
    assert oracle.crc32(b"") == 0


def test_crc_matches_reference_source():
    so = os.path.join(ROOT, "oracle", "_ref", "libref_crc32.so")
    if not os.path.exists(so):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    ref = ctypes.CDLL(so)
    f = getattr(ref, "_Z11bzip2_crc32PKhm")
    f.restype = ctypes.c_uint32
    f.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    rng = np.random.default_rng(1)
    for n in [0, 1, 7, 900000]:
        d = rng.integers(0, 256, n, dtype=np.uint8)
        assert f(d.ctypes.data_as(ctypes.c_void_p), n) == oracle.crc32(d)


# ----------------------------------------------------------------- RLE1
@pytest.mark.parametrize("max_run", [20, 3, 10000])
def test_rle1_reference_inputs(max_run):
    data = refinputs.rle1_input(max_run)
    got = oracle.rle1(data).tobytes()
    assert got == spec.rle1(data)
    if max_run == 3:  # rle1_incompressible: output equals input
        assert got == data


def test_rle1_closed_form_runs():
    for c in [1, 2, 3, 4, 5, 6, 258, 259, 260, 262, 263, 264, 518, 1000, 900000]:
        got = oracle.rle1(b"\x07" * c)
        full, rem = divmod(c, 259)
        assert got.size == 5 * full + (rem if rem < 4 else 5)
        assert got.tobytes() == spec.rle1(b"\x07" * c)


# ------------------------------------------------------------------ BWT
@pytest.mark.parametrize("alphabet", [256, 4, 1])
def test_bwt_reference_inputs(alphabet):
    data = refinputs.bwt_input(alphabet)
    sa, _ = oracle.bwt(data)
    np.testing.assert_array_equal(sa, spec.bwt_bruteforce(data))
    if alphabet == 1:  # fbwt_single_character: identical rotations -> identity
        np.testing.assert_array_equal(sa, np.arange(5000))


@pytest.mark.parametrize("seed", range(6))
def test_bwt_bruteforce_small(seed):
    rng = np.random.default_rng(seed)
    for n in [1, 2, 3, 5, 17, 64, 301]:
        a = int(rng.integers(1, 5))
        if seed % 2:
            base = rng.integers(0, a + 1, int(rng.integers(1, 6)), dtype=np.uint8).tobytes()
            data = (base * (n // len(base) + 1))[:n]
        else:
            data = rng.integers(0, a + 1, n, dtype=np.uint8).tobytes()
        sa, _ = oracle.bwt(data)
        np.testing.assert_array_equal(sa, spec.bwt_bruteforce(data))


def test_bwt_inverse_roundtrip():
    rng = np.random.default_rng(3)
    data = rng.integers(0, 3, 20000, dtype=np.uint8).tobytes() + b"x"
    sa, _ = oracle.bwt(data)
    last = bytes(data[(i - 1) % len(data)] for i in sa)
    op = int(np.nonzero(sa == 0)[0][0])
    assert spec.inverse_bwt(last, op) == data


# ------------------------------------------------------------------ MTF
def test_symbols_table_reference_case():
    table, data = refinputs.symbols_table_case()
    present = sorted(set(data))
    assert present == table


@pytest.mark.parametrize("alphabet", [256, 4])
def test_mtf_reference_inputs(alphabet):
    data = refinputs.mtf_input(alphabet)
    n = len(data)
    sa = (np.arange(n, dtype=np.int64) + 1) % n  # L == input
    out, op, nuse, pres = oracle.mtf(data, sa.astype(np.int32))
    assert nuse == len(set(data))
    assert op == n - 1
    ref = spec.mtf(data[:200000])
    np.testing.assert_array_equal(out[:200000], ref)
    np.testing.assert_array_equal(oracle.mtf_plain(data), out)


# ----------------------------------------------------------------- RLE2
@pytest.mark.parametrize("case", ["r20", "incompressible", "r10000"])
def test_rle2_reference_inputs(case):
    data = {"r20": lambda: refinputs.rle2_input(20), "r10000": lambda: refinputs.rle2_input(10000),
            "incompressible": refinputs.rle2_incompressible_input}[case]()
    ts = refinputs.table_size(data)
    got = oracle.rle2(data, ts)
    np.testing.assert_array_equal(got, spec.rle2(data, ts))


def test_rle2_zero_run_digits():
    # a zero run of k is the bits of k+1 below its leading one, LSB first
    for k in range(1, 300):
        got = oracle.rle2(b"\x00" * k, 1)[:-1]
        v = k + 1
        want = [(v >> i) & 1 for i in range(v.bit_length() - 1)]
        assert got.tolist() == want


# -------------------------------------------------------------- Huffman
def test_code_lengths_invariants():
    rng = np.random.default_rng(5)
    for alpha in [3, 10, 50, 258]:
        for _ in range(5):
            f = (rng.pareto(1.0, alpha) * 100).astype(np.int32)
            ln = oracle.make_code_lengths(f, alpha, 17)
            assert ln.max() <= 17 and ln.min() >= 1
            assert sum(2.0 ** -int(x) for x in ln) <= 1.0 + 1e-12


def _tree_height_exceeds(freq, max_len=17):
    """Height of the libbz2 tree from the weight values alone: leaves sorted
    by (weight, id), then a two-queue merge (merged weights come out in
    non-decreasing order), depth byte = 1 + the deeper input's."""
    w = sorted(((max(int(f), 1) << 8), i) for i, f in enumerate(freq))
    leaves = [x for x, _ in w]
    merged, i1, i2 = [], 0, 0
    for _ in range(len(leaves) - 1):
        ab = []
        for _t in range(2):
            if i1 < len(leaves) and (i2 >= len(merged) or leaves[i1] <= merged[i2]):
                ab.append(leaves[i1]); i1 += 1
            else:
                ab.append(merged[i2]); i2 += 1
        d = 1 + max(ab[0] & 0xFF, ab[1] & 0xFF)
        if d > max_len:
            return True
        merged.append(((ab[0] & ~0xFF) + (ab[1] & ~0xFF)) | d)
    return False


def test_halving_rounds_follow_from_weight_values():
    # The device decides libbz2's weight-halving rounds (huffman.cu:110-123)
    # from the tree height alone and then builds once: one build with the
    # final weights must equal the retrying build.
    rng = np.random.default_rng(11)
    fib = [1, 1]
    while len(fib) < 30:  # Fib(30) < 900000, the largest count a block can hold
        fib.append(fib[-1] + fib[-2])
    cases = []
    for alpha in [3, 20, 30, 64, 130, 258]:
        cases.append(np.array(fib[:alpha] if alpha <= 30 else fib + [1] * (alpha - 30)))
        cases.append((rng.pareto(0.5, alpha) * 50).astype(np.int64) % 900000)
        cases.append(np.where(rng.random(alpha) < 0.7, 0, rng.integers(1, 900000, alpha)))
    rounds_seen = 0
    for f in cases:
        if f.sum() > 900000:  # a block holds at most 900000 + 1 symbols
            f = np.maximum(f * 900000 // f.sum(), 0)
        alpha = f.size
        want = oracle.make_code_lengths(f.astype(np.int32), alpha, 17)
        g = np.maximum(f, 1)
        rounds = 0
        while _tree_height_exceeds(g):
            g = 1 + g // 2
            rounds += 1
        rounds_seen += rounds
        got = oracle.make_code_lengths(g.astype(np.int32), alpha, 64)  # no further rounds
        np.testing.assert_array_equal(got, want)
        assert got.max() <= 17
    assert rounds_seen > 0


def test_huff_reference_encoding_roundtrip():
    data = refinputs.encoding_input()
    ln, cd, sel, ng = oracle.huff_build(data, 15)
    assert sel.size == 200 and sel.max() < ng
    words, bits = oracle.huff_encode(data, ln, cd, sel)
    dec = spec.decode_huffman(words, bits, ln[:ng], cd[:ng], sel, 17)
    np.testing.assert_array_equal(dec, data)


def test_huff_stability_selectors():
    for a, data in refinputs.stability_inputs():
        if a % 37 != 1:
            continue
        ln, cd, sel, ng = oracle.huff_build(data, a)
        assert ng == 6 and sel.max() <= 5


def test_selector_mtf_spec():
    rng = np.random.default_rng(2)
    sel = rng.integers(0, 6, 1000).astype(np.uint8)
    got = oracle.selector_mtf(sel, 6)
    lst = list(range(6))
    for i, v in enumerate(sel):
        j = lst.index(int(v))
        assert got[i] == j
        lst.insert(0, lst.pop(j))


# ---------------------------------------------------- whole stream / libbz2
def test_empty_stream_known_answer():
    # tests/compression/compress.cu:28-50
    assert oracle.compress(b"", 9) == bytes.fromhex("425a683917724538509000000000")
    assert bz2.compress(b"", 9) == oracle.compress(b"", 9)


def _stock_equal_case(data: bytes, level: int):
    # stock-equivalent blocks: RLE1 length <= 100000*level - 19, no run > 255,
    # no identical rotations (checked by the presence of a unique byte)
    r1 = oracle.rle1(data)
    assert r1.size <= 100000 * level - 19
    return oracle.compress(data, level, block_bytes=len(data)) == bz2.compress(data, level)


@pytest.mark.parametrize("seed", range(4))
def test_libbz2_byte_equality(seed):
    import bzip2gpu_amd.datagen as dg
    n = [1000, 50000, 99000, 880000][seed]
    level = [1, 3, 1, 9][seed]
    data = dg.markov(n, 100 + seed, order=1 + seed % 2, s=1.1).tobytes()
    assert _stock_equal_case(data, level)


def test_libbz2_byte_equality_random_and_runs():
    rng = np.random.default_rng(9)
    data = rng.integers(0, 256, 899000, dtype=np.uint8).tobytes()
    assert _stock_equal_case(data, 9)
    runs = b"".join(bytes([int(v)]) * int(l) for v, l in zip(rng.integers(0, 256, 3000), rng.integers(1, 200, 3000)))
    assert _stock_equal_case(runs[:99000], 1)


def _deepest_unlimited_table(data: bytes) -> int:
    """Depth of the deepest Huffman tree of the block's last iteration if
    there were no 17-bit limit: the tables' symbol counts from the final
    selectors, built with a limit no tree reaches."""
    r1 = np.frombuffer(oracle.rle1(data), np.uint8)
    sa, _ = oracle.bwt(r1)
    mtfb, _, nuse = oracle.mtf(r1.tobytes(), sa)[:3]
    sym = oracle.rle2(mtfb, nuse)
    _, _, sel, ng = oracle.huff_build(sym, nuse)
    alpha = nuse + 2
    deepest = 0
    for t in range(ng):
        f = np.zeros(alpha, np.int32)
        for g, tb in enumerate(sel):
            if tb == t:
                np.add.at(f, sym[g * 50:(g + 1) * 50], 1)
        deepest = max(deepest, int(oracle.make_code_lengths(f, alpha, 255).max()))
    return deepest


def test_libbz2_byte_equality_with_weight_halving():
    # i.i.d. bytes with p(b) proportional to 0.7^b: one of the last
    # iteration's tables would be 18 levels deep, so libbz2 halves the weights
    # and rebuilds it (huffman.cu:110-123); the stream still equals stock
    # libbz2's, which pins that branch of the oracle
    rng = np.random.default_rng(4)
    p = 0.7 ** np.arange(256)
    data = rng.choice(256, size=880000, p=p / p.sum()).astype(np.uint8).tobytes()
    assert _deepest_unlimited_table(data) > 17
    assert _stock_equal_case(data, 9)


def test_multiblock_roundtrip_and_combined_crc():
    import bzip2gpu_amd.datagen as dg
    from tests import bz2dec
    data = dg.workload_text1().tobytes()
    out = oracle.compress(data, 9)
    dec, sizes = bz2dec.decompress(out)
    assert dec == data
    # the first 900000-byte block grows past 900000 bytes under RLE1, so
    # stock libbz2 refuses the stream for its size alone
    assert sizes[0] > 900000
    with pytest.raises(OSError):
        bz2.decompress(out)
    # combined CRC closed form: C = XOR_i rotl(crc_i, (N-1-i) mod 32)
    crcs = [oracle.crc32(data[i:i + 900000]) for i in range(0, len(data), 900000)]
    c = 0
    for i, v in enumerate(crcs):
        r = (len(crcs) - 1 - i) % 32
        c ^= (((v << r) | (v >> (32 - r))) & 0xFFFFFFFF) if r else v
    assert int.from_bytes(_trailer_crc(out), "big") == c


def _trailer_crc(out: bytes) -> bytes:
    # the stream CRC is the 32 bits after the end-of-stream magic, before
    # the byte padding
    bits = np.unpackbits(np.frombuffer(out, np.uint8))
    magic = np.unpackbits(np.frombuffer(bytes.fromhex("177245385090"), np.uint8))
    for pad in range(8):
        end = bits.size - pad
        if np.array_equal(bits[end - 80:end - 32], magic):
            return np.packbits(bits[end - 32:end]).tobytes()
    raise AssertionError("no end-of-stream magic")


def test_decoder_agrees_with_libbz2():
    from tests import bz2dec
    import bzip2gpu_amd.datagen as dg
    data = dg.mix(300000, 4).tobytes()
    stock = bz2.compress(data, 1)
    assert bz2dec.decompress(stock)[0] == data


def test_periodic_roundtrip():
    data = (b"abcab" * 40000)[:199999]
    out = oracle.compress(data, 2, block_bytes=100000)
    assert bz2.decompress(out) == data


def test_reference_cut_levels():
    # the reference cuts 900000-byte blocks at every level; the level only
    # changes the header digit
    import bzip2gpu_amd.datagen as dg
    data = dg.markov(300000, 7, order=1).tobytes()
    a = oracle.compress(data, 9)
    b = oracle.compress(data, 3)
    assert a[3:4] == b"9" and b[3:4] == b"3" and a[4:] == b[4:]
