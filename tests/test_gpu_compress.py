"""Whole-path parity: .bz2 streams from the batched and plain HIP paths are
byte-identical to the oracle's and decode back to the input."""
import bz2
import sys

import numpy as np
import pytest

import oracle
from tests import bz2dec

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _same(a: bytes, b: bytes, what: str) -> None:
    """Byte equality without pytest's (very slow) diff of large bytes."""
    if a == b:
        return
    n = min(len(a), len(b))
    first = next((i for i in range(n) if a[i] != b[i]), n)
    raise AssertionError(f"{what}: lengths {len(a)} vs {len(b)}, first difference at byte {first}")


def _inputs():
    import bzip2gpu_amd.datagen as dg
    rng = np.random.default_rng(11)
    return [
        ("empty", b""),
        ("one", b"x"),
        ("hello", b"hello world!\n"),
        ("words", b"".join(b"word%d\n" % i for i in range(10))),
        ("zeros900k", b"\x00" * 900000),
        ("text1_1MiB", dg.workload_text1().tobytes()),
        ("random_899k", rng.integers(0, 256, 899000, dtype=np.uint8).tobytes()),
        ("mix_2MiB", dg.mix(2 << 20, 5).tobytes()),
        ("two_blocks_exact", dg.markov(1800000, 6, order=2, s=1.2).tobytes()),
        ("periodic", dg.periodic(950000, 8, 7, 0.001).tobytes()),
    ]


@pytest.mark.parametrize("path", ["batched", "plain"])
def test_compress_matches_oracle(dev, path):
    import bzip2gpu_amd as bz
    fn = bz.compress if path == "batched" else bz.compress_plain
    for name, data in _inputs():
        _case(f"compress {path} {name} len={len(data)}")
        got = fn(data, 9)
        want = oracle.compress(data, 9)
        _same(got, want, f"{path} {name}")


def test_levels_and_block_sizes(dev):
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = dg.mix(700000, 12).tobytes()
    for level in [1, 3, 9]:
        for bb in [900000, 100000 * level]:
            _case(f"level={level} block_bytes={bb}")
            got = bz.compress(data, level, block_bytes=bb)
            _same(got, oracle.compress(data, level, block_bytes=bb), f"level {level} bb {bb}")
            dec, sizes = bz2dec.decompress(got)
            assert dec == data
            if max(sizes) <= 100000 * level:
                assert bz2.decompress(got) == data


def test_round0_tile_edges(dev):
    """Round 0's first radix pass makes keys from the text: blocks just above
    the small-group limit (1024 rows) and blocks whose last 4096-row tile has
    1-3 rows (the key and the digit wrap around the block end).  Random bytes
    keep RLE1 from changing the lengths."""
    import bzip2gpu_amd as bz
    rng = np.random.default_rng(21)
    for n in [1025, 1026, 1027, 1030, 4095, 4096, 4097, 4098, 4099, 8193, 12290, 40963]:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        _case(f"round-0 tile edges n={n}")
        _same(bz.compress(data, 9), oracle.compress(data, 9), f"n={n}")
    # a text-like block with long equal-key groups across tile edges
    data = (b"abcabcabd" * 1000)[:8195]
    _case("round-0 tile edges periodic n=8195")
    _same(bz.compress(data, 9), oracle.compress(data, 9), "periodic 8195")


def test_stock_equal_blocks(dev):
    import bzip2gpu_amd as bz
    rng = np.random.default_rng(2)
    data = rng.integers(0, 256, 899000, dtype=np.uint8).tobytes()
    _case("stock-equal random block")
    _same(bz.compress(data, 9), bz2.compress(data, 9), "stock-equal")
    # a block whose last-iteration tables need libbz2's weight halving
    # (bytes with p(b) proportional to 0.7^b; tests/test_oracle.py checks the
    # unlimited tree depth of this input is 18)
    p = 0.7 ** np.arange(256)
    data = np.random.default_rng(4).choice(256, size=880000, p=p / p.sum()).astype(np.uint8).tobytes()
    _case("stock-equal block with weight halving")
    _same(bz.compress(data, 9), bz2.compress(data, 9), "stock-equal, halving")


def test_batch_device_many_inputs(dev):
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    rng = np.random.default_rng(3)
    sizes = [0, 1, 1000, 4096, 100000, 250001, 900000, 1234567]
    datas = [dg.mix(s, 40 + i).tobytes() if s else b"" for i, s in enumerate(sizes)]
    levels = rng.integers(1, 10, len(sizes)).astype(np.int32)
    offs = np.zeros(len(sizes), np.uint64)
    buf = bytearray()
    for i, d in enumerate(datas):
        offs[i] = len(buf)
        buf += d
        buf += b"\x00" * ((-len(buf)) % 16)
    _case(f"batch device {len(sizes)} inputs")
    t = torch.from_numpy(np.frombuffer(bytes(buf) + b"\x00" * 16, np.uint8).copy()).to(dev)
    ptr, bo, by = bz.compress_batch_device(t, offs, [len(d) for d in datas], levels, 0)
    import ctypes
    total = int(bo[-1] + by[-1])
    host = (ctypes.c_uint8 * total)()
    torch.cuda.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipMemcpy(host, ctypes.c_void_p(ptr), ctypes.c_size_t(total), 2) == 0
    host = bytes(host)
    for i, d in enumerate(datas):
        s = host[bo[i]:bo[i] + by[i]]
        _same(s, oracle.compress(d, int(levels[i]), block_bytes=100000 * int(levels[i])), f"input {i}")


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_radix_shapes_identical(dev, variant):
    """Every autotuned launch shape of the radix scatter gives the same stream."""
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = dg.mix(3 << 20, 77).tobytes()
    want = oracle.compress(data, 9)
    try:
        bz.tune_force(0, variant)
        _case(f"radix shape variant {variant}")
        _same(bz.compress(data, 9), want, f"variant {variant}")
    finally:
        bz.tune_force(0, -1)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_mtf_segment_runs_identical(dev, variant):
    """Every MTF run length (1, 2, 4 or 8 segments per wave) gives the
    oracle's stream, with runs that cross block ends (100000-byte blocks
    span 25 tiles) and a block shorter than a run."""
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = dg.mix(3 << 20, 81).tobytes()
    try:
        bz.tune_force(2, variant)
        for level in (9, 1):
            _case(f"mtf segment run variant {variant}, level {level}")
            bb = 100000 * level
            _same(bz.compress(data, level, block_bytes=bb), oracle.compress(data, level, block_bytes=bb),
                  f"variant {variant} level {level}")
        short = data[:30000]
        _same(bz.compress(short, 9), oracle.compress(short, 9), f"variant {variant} short")
    finally:
        bz.tune_force(2, -1)


@pytest.mark.parametrize("policy", [0, 1])
def test_rank_building_paths(dev, policy):
    """Blocks whose suffix-sort ranks are never built (text-like, groups of
    at most 64 rows after round 0), built in round 1 (groups of about 1000
    rows from a 1000-byte period), kept from round 0 (tile-long groups from
    a 2-byte period), built in round 3 or later (a 65536-byte period with
    mutations), and small blocks with long groups, in one batch: the oracle's
    stream with both launch policies."""
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    parts = [dg.mix(1 << 20, 82), dg.periodic(900000, 83, 1000, 0.001), dg.periodic(900000, 84, 2, 0.001),
             dg.periodic(900000, 85, 65536, 0.001), np.zeros(3000, np.uint8), dg.periodic(1000, 86, 3, 0.0),
             dg.mix(700000, 87)]
    data = np.concatenate(parts).tobytes()
    want = oracle.compress(data, 9, threads=8)
    try:
        bz.tune_force(1, policy)
        _case(f"rank building paths, sort policy {policy}")
        _same(bz.compress(data, 9), want, f"policy {policy}")
    finally:
        bz.tune_force(1, -1)


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_sort_launch_policies_identical(dev, policy):
    """Both launch policies of the suffix-sort rounds (per round with exact
    grids; device-driven persistent grids) give the oracle's stream, on a
    batch that mixes short blocks, full blocks and a periodic block that
    needs many rounds, and report the oracle's round count."""
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    per = dg.periodic(900000, 31, 7000, 0.001)
    data = np.concatenate([dg.mix(3 << 20, 79), per])
    want = oracle.compress(data, 9, threads=8)
    try:
        bz.tune_force(0, 1)
        bz.tune_force(1, policy)
        for graphs in (False, True, True):
            _case(f"sort launch policy {policy}, stage graphs {graphs}")
            bz.set_stage_graphs(graphs)
            _same(bz.compress(data.tobytes(), 9), want, f"policy {policy} graphs {graphs}")
            torch.cuda.synchronize()
        _, rounds, _ = bz.last_stage_times()
        r1 = np.frombuffer(oracle.rle1(per.tobytes()), np.uint8)
        assert rounds >= oracle.bwt(r1)[1] - 1
    finally:
        bz.set_stage_graphs(True)
        bz.tune_force(1, -1)
        bz.tune_force(0, -1)


def test_autotuner_explores_then_chooses(dev):
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = dg.mix(2 << 20, 78).tobytes()
    want = oracle.compress(data, 9)
    _case("autotuner exploration")
    for _ in range(8):  # 3 shapes x 2 samples for the whole-block class
        _same(bz.compress(data, 9), want, "autotuned run")
    chosen, rows = bz.tune_state(0, 0)
    assert [v for v, _, _ in rows] == [256, 512, 1024]
    assert chosen in (0, 1, 2)
    assert all(n >= 2 for _, _, n in rows)


@pytest.mark.parametrize("policy", [0, 1])
def test_stage_graph_replays_match_eager(dev, policy):
    """Stage groups captured as HIP graphs (with the device-driven sort
    policy, the suffix-sort rounds too): every replay, with eager work and
    synchronisations between replays, equals the eager result."""
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    a = dg.mix(2 << 20, 90).tobytes()
    b = dg.mix(1 << 20, 91).tobytes()
    _case(f"stage graphs, sort policy {policy}: eager references")
    bz.tune_force(0, 0)
    bz.tune_force(1, policy)
    bz.set_stage_graphs(False)
    eager_a, eager_b = bz.compress(a, 9), bz.compress(b, 9)
    _same(eager_a, oracle.compress(a, 9), "eager a")
    try:
        bz.set_stage_graphs(True)
        _same(bz.compress(a, 9), eager_a, "first graph run")  # captures
        c0, r0 = bz.graph_counts()
        for i in range(3):
            _case(f"stage graphs: replay {i}")
            _same(bz.compress(a, 9), eager_a, f"replay {i}")
            torch.cuda.synchronize()
            bz.set_stage_graphs(False)
            _same(bz.compress(b, 9), eager_b, f"eager between replays {i}")
            bz.set_stage_graphs(True)
        c1, r1 = bz.graph_counts()
        assert c1 == c0, "replays of an unchanged plan must not re-capture"
        # five stage groups per run, plus round 0 and the later rounds of
        # the sort when the rounds are device-driven
        assert r1 - r0 >= 3 * (5 if policy == 0 else 7)
    finally:
        bz.set_stage_graphs(True)
        bz.tune_force(1, -1)
        bz.tune_force(0, -1)


def test_plain_threads_matches_batched(dev):
    """The reference-style plain path (host threads, one HIP stream and one
    pipeline each, one block per run) gives the batched path's bytes."""
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    sizes = [0, 1, 5000, 250000, 1234567, 2 * 900000 + 3, 100000]
    levels = np.array([9, 1, 4, 2, 9, 9, 1], np.int32)
    offs = np.cumsum([0] + [s + (-s) % 64 for s in sizes[:-1]]).astype(np.uint64)
    pool = np.zeros(int(offs[-1]) + sizes[-1] + 64, np.uint8)
    for i, sz in enumerate(sizes):
        pool[int(offs[i]):int(offs[i]) + sz] = dg.mix(sz, 60 + i) if sz else 0
    t = torch.from_numpy(pool).to(dev)
    _case(f"plain threads: {len(sizes)} inputs, 4 threads")
    got = bz.compress_plain_threads(t, offs, sizes, levels, 0, 4)
    for i, sz in enumerate(sizes):
        lv = int(levels[i])
        _same(got[i], oracle.compress(pool[int(offs[i]):int(offs[i]) + sz], lv, block_bytes=100000 * lv),
              f"input {i}")
