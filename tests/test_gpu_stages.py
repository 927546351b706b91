"""Per-stage parity of the HIP kernels (through the C ABI) with the oracle
and the reference's own test specs, on the reference's seeded inputs."""
import sys

import numpy as np
import pytest

import oracle
from tests import refinputs, spec

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _t(dev, data, dtype=None):
    import torch
    a = np.frombuffer(data, np.uint8) if isinstance(data, (bytes, bytearray)) else np.asarray(data)
    t = torch.from_numpy(a.copy())
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)


def test_crc32(dev):
    import bzip2gpu_amd as bz
    rng = np.random.default_rng(0)
    for n in [1, 5, 4095, 4096, 4097, 65536, 900000, 1000000]:
        _case(f"crc32 n={n}")
        d = rng.integers(0, 256, n, dtype=np.uint8)
        assert bz.crc32(_t(dev, d)) == oracle.crc32(d)
    assert bz.crc32(_t(dev, b"123456789")) == 0xFC891918


@pytest.mark.parametrize("max_run", [20, 3, 10000])
def test_rle1_reference(dev, max_run):
    import bzip2gpu_amd as bz
    _case(f"rle1 RUN_RLE1({max_run})")
    data = refinputs.rle1_input(max_run)
    got = bz.rle1(_t(dev, data)).cpu().numpy()
    np.testing.assert_array_equal(got, oracle.rle1(data))


def test_rle1_edges(dev):
    import bzip2gpu_amd as bz
    cases = [b"a", b"aaaa", b"aaaaa", b"\x00" * 259, b"\x00" * 260, b"\x05" * 900000, b"\x05" * 4096 + b"\x06" * 4100,
             (b"ab" * 3000), bytes(range(256)) * 100]
    rng = np.random.default_rng(5)
    runs = b"".join(bytes([int(v)]) * int(l) for v, l in zip(rng.integers(0, 3, 4000), rng.integers(1, 600, 4000)))
    cases.append(runs[:900000])
    for c in cases:
        _case(f"rle1 edge len={len(c)}")
        got = bz.rle1(_t(dev, c)).cpu().numpy()
        np.testing.assert_array_equal(got, oracle.rle1(c))


@pytest.mark.parametrize("alphabet", [256, 4, 1])
def test_bwt_reference(dev, alphabet):
    import bzip2gpu_amd as bz
    _case(f"bwt RUN_FBWT({alphabet})")
    data = refinputs.bwt_input(alphabet)
    sa, rounds = bz.bwt(_t(dev, data))
    np.testing.assert_array_equal(sa.cpu().numpy(), spec.bwt_bruteforce(data))


@pytest.mark.parametrize("kind", ["random", "text", "runs", "period2", "period7", "period1000", "tiny"])
def test_bwt_vs_oracle(dev, kind):
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    n = 300000
    if kind == "random":
        data = dg.uniform(n, 1)
    elif kind == "text":
        data = dg.markov(n, 2, order=2, s=1.2)
    elif kind == "runs":
        data = oracle.rle1(dg.runs(n, 3))
    elif kind.startswith("period"):
        data = dg.periodic(n, 4, int(kind[6:]), 0.001)
    else:
        for m in [1, 2, 3, 4, 5, 7, 64, 1023, 1024, 1025, 2047, 2048, 2049, 5000]:
            _case(f"bwt tiny m={m}")
            d = dg.markov(m, m, order=1, s=2.0)
            sa, _ = bz.bwt(_t(dev, d))
            np.testing.assert_array_equal(sa.cpu().numpy(), oracle.bwt(d)[0])
        return
    _case(f"bwt {kind} m={data.size}")
    sa, rounds = bz.bwt(_t(dev, data))
    ref, _ = oracle.bwt(data)
    np.testing.assert_array_equal(sa.cpu().numpy(), ref)


def test_bwt_identical_rotations(dev):
    import bzip2gpu_amd as bz
    for base, reps in [(b"a", 5000), (b"ab", 3000), (b"abcab", 700), (bytes(range(7)), 20000)]:
        d = base * reps
        _case(f"bwt periodic base={len(base)} m={len(d)}")
        sa, _ = bz.bwt(_t(dev, d))
        np.testing.assert_array_equal(sa.cpu().numpy(), oracle.bwt(d)[0])


def test_bwt_rank_state_between_calls(dev):
    """Consecutive sorts on one engine: a block that kept its ranks (a
    tile-long group in round 0) must not leave that state to the next sort,
    whose ranks are then built only where that sort needs them."""
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    seq = [b"a" * 20000, dg.periodic(200000, 4, 64, 0.0005), bytes(range(7)) * 9000,
           dg.markov(250000, 5, order=2, s=1.2), b"a" * 20000]
    for i, d in enumerate(seq):
        d = np.frombuffer(d, dtype=np.uint8) if isinstance(d, bytes) else d
        _case(f"bwt sequence step={i} m={d.size}")
        sa, _ = bz.bwt(_t(dev, d))
        np.testing.assert_array_equal(sa.cpu().numpy(), oracle.bwt(d)[0])


def test_symbols_table(dev):
    import bzip2gpu_amd as bz
    _case("symbols table")
    table, data = refinputs.symbols_table_case()
    np.testing.assert_array_equal(bz.symbols_table(_t(dev, data)), np.array(table, np.uint8))


@pytest.mark.parametrize("alphabet", [256, 4])
def test_mtf_reference(dev, alphabet):
    import torch
    import bzip2gpu_amd as bz
    _case(f"mtf RUN_FMTF({alphabet})")
    data = refinputs.mtf_input(alphabet)
    n = len(data)
    sa = ((np.arange(n) + 1) % n).astype(np.int32)
    out, op, nuse, pres = bz.mtf(_t(dev, data), torch.from_numpy(sa).to(dev))
    ref, rop, rnuse, rpres = oracle.mtf(data, sa)
    assert (op, nuse) == (rop, rnuse)
    np.testing.assert_array_equal(pres, rpres)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)


def test_mtf_late_first_occurrence(dev):
    """A present byte whose first occurrence is past the first 4096-symbol
    segment, with most byte values absent from the block."""
    import torch
    import bzip2gpu_amd as bz
    rng = np.random.default_rng(21)
    L = rng.integers(10, 14, 30000).astype(np.uint8)
    L[20000] = 200
    L[25000] = 3
    n = L.size
    sa = ((np.arange(n) + 1) % n).astype(np.int32)
    _case("mtf late first occurrence")
    out, op, nuse, _ = bz.mtf(_t(dev, L), torch.from_numpy(sa).to(dev))
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.mtf_plain(L))


def test_mtf_after_bwt(dev):
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    for kind, data in [("text", dg.markov(200000, 9, order=2, s=1.2)), ("mix", oracle.rle1(dg.mix(200000, 3)))]:
        _case(f"mtf after bwt {kind}")
        d = _t(dev, data)
        sa, _ = bz.bwt(d)
        out, op, nuse, _ = bz.mtf(d, sa)
        ref, rop, rnuse, _ = oracle.mtf(data, oracle.bwt(data)[0])
        assert (op, nuse) == (rop, rnuse)
        np.testing.assert_array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("case", ["r20", "incompressible", "r10000"])
def test_rle2_reference(dev, case):
    import bzip2gpu_amd as bz
    _case(f"rle2 {case}")
    data = {"r20": lambda: refinputs.rle2_input(20), "r10000": lambda: refinputs.rle2_input(10000),
            "incompressible": refinputs.rle2_incompressible_input}[case]()
    ts = refinputs.table_size(data)
    got = bz.rle2(_t(dev, data), ts).cpu().numpy().astype(np.uint16)
    np.testing.assert_array_equal(got, oracle.rle2(data, ts))


def test_rle2_edges(dev):
    import bzip2gpu_amd as bz
    for c in [b"\x00", b"\x00" * 4096, b"\x00" * 5000 + b"\x01", b"\x01" + b"\x00" * 9000, b"\x03\x00\x00\x02" * 3000]:
        _case(f"rle2 edge len={len(c)}")
        got = bz.rle2(_t(dev, c), 4).cpu().numpy().astype(np.uint16)
        np.testing.assert_array_equal(got, oracle.rle2(c, 4))


def test_histogram_reference(dev):
    import torch
    import bzip2gpu_amd as bz
    _case("histogram")
    data = refinputs.histogram_input()
    f = bz.histogram(torch.from_numpy(data.astype(np.int16)).to(dev)).cpu().numpy()
    np.testing.assert_array_equal(f, np.bincount(data, minlength=258)[:258])


def test_huff_build_stability_and_parity(dev):
    import torch
    import bzip2gpu_amd as bz
    for a, data in refinputs.stability_inputs():
        if a % 16 != 1 and a != 255:
            continue
        _case(f"huff build alphabet={a}")
        ln, cd, sel, ng = bz.huff_build(torch.from_numpy(data.astype(np.int16)).to(dev), a)
        rln, rcd, rsel, rng_ = oracle.huff_build(data, a)
        assert ng == rng_
        assert sel.max() <= 5
        np.testing.assert_array_equal(sel, rsel)
        np.testing.assert_array_equal(ln[:ng, :a + 2], rln[:ng, :a + 2])
        np.testing.assert_array_equal(cd[:ng, :a + 2], rcd[:ng, :a + 2])


def test_huff_build_halving_rounds(dev):
    # skewed symbol counts whose trees exceed 17 levels, so libbz2 halves the
    # weights (huffman.cu:110-123): lengths, codes and selectors must still
    # equal the oracle's
    import torch
    import bzip2gpu_amd as bz
    rng = np.random.default_rng(21)
    for a in [24, 254]:
        _case(f"huff build halving alphabet={a}")
        # symbol k about 450000 / 2^k times: the main table's tree is 20 deep
        counts = [max(1, 450000 >> k) for k in range(a + 1)] + [1]
        data = np.repeat(np.arange(a + 2, dtype=np.uint16), counts)
        data = data[data != a + 1]
        rng.shuffle(data)
        data = np.concatenate([data, np.array([a + 1], np.uint16)])
        ln, cd, sel, ng = bz.huff_build(torch.from_numpy(data.astype(np.int16)).to(dev), a)
        rln, rcd, rsel, rng_ = oracle.huff_build(data, a)
        assert ng == rng_
        np.testing.assert_array_equal(sel, rsel)
        np.testing.assert_array_equal(ln[:ng, :a + 2], rln[:ng, :a + 2])
        np.testing.assert_array_equal(cd[:ng, :a + 2], rcd[:ng, :a + 2])
        assert rln[:ng, :a + 2].max() <= 17


def test_huff_encoding_roundtrip(dev):
    import torch
    import bzip2gpu_amd as bz
    _case("huff encode")
    data = refinputs.encoding_input()
    d = torch.from_numpy(data.astype(np.int16)).to(dev)
    ln, cd, sel, ng = bz.huff_build(d, 15)
    words, bits = bz.huff_encode(d, ln, cd, sel, ng)
    rwords, rbits = oracle.huff_encode(data, ln, cd, sel)
    assert bits == rbits
    np.testing.assert_array_equal(words.cpu().numpy().view(np.uint32), rwords)
    dec = spec.decode_huffman(words.cpu().numpy().view(np.uint32), bits, ln[:ng], cd[:ng], sel, 17)
    np.testing.assert_array_equal(dec, data)
