"""Parity at the shapes of the benchmark workloads, against the oracle.

* Adversarial suffix sort: whole 900000-byte blocks that repeat a seeded
  random string of period 1, 2, 7, 1000 and 65536 with 0.1 % single-byte
  mutations (the periodic workload's own generator), all in one stream.
* Batched uneven inputs: sizes up to 8 MiB at levels 1..9 with
  100000*level blocks, run with a small sub-batch cap so that the inputs
  cross at least three device sub-batches.
* Serving at the serving workload's 64 KiB median size, and live
  submit/wait requests mixed with an open-loop replay.
"""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _same(a: bytes, b: bytes, what: str) -> None:
    if a == b:
        return
    n = min(len(a), len(b))
    first = next((i for i in range(n) if a[i] != b[i]), n)
    raise AssertionError(f"{what}: lengths {len(a)} vs {len(b)}, first difference at byte {first}")


def _d2h(ptr, nbytes):
    import ctypes
    import torch
    torch.cuda.synchronize()
    host = (ctypes.c_uint8 * max(1, nbytes))()
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipMemcpy(host, ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 2) == 0
    return bytes(host)[:nbytes]


def test_periodic_full_blocks(dev):
    """Five full blocks, periods 1/2/7/1000/65536 in block order, as the
    periodic workload lays them out; the whole stream (origPtr, last column,
    tables, payload) equals the oracle's, and the rounds reached match."""
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = dg.workload_periodic(5 * 900000)
    _case(f"periodic workload, 5 blocks, {data.size} bytes")
    got = bz.compress(data.tobytes(), 9)
    _, rounds, _ = bz.last_stage_times()
    want = oracle.compress(data, 9, threads=5)
    _same(got, want, "periodic 5 blocks")
    # The oracle doubles from 1-byte classes and stops at the first k with
    # 2^k > L (L = the block's longest common prefix of two rotations); the
    # device sort starts from 4-byte keys and runs a round while groups
    # remain, floor(log2 L) rounds for L >= 4: one fewer.
    k = max(oracle.bwt(oracle.rle1(data[j * 900000:(j + 1) * 900000].tobytes()))[1] for j in range(5))
    assert rounds == k - 1, (rounds, k)


@pytest.mark.parametrize("period", [1000, 65536])
def test_periodic_block_sa(dev, period):
    """The suffix array of a mutated periodic block's RLE1 text, exactly."""
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    blk = dg.periodic(900000, 4242 + period, period, 0.001)
    r1 = np.frombuffer(oracle.rle1(blk.tobytes()), np.uint8)
    _case(f"periodic block SA period={period} m={r1.size}")
    d = torch.from_numpy(r1.copy()).to(dev)
    sa, rounds = bz.bwt(d)
    want, wrounds = oracle.bwt(r1)
    assert (sa.cpu().numpy() == want).all(), f"period {period}: suffix array differs"
    assert rounds >= 2


def test_batch_inputs_up_to_8mib_across_sub_batches(dev):
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    sizes = [8 * MiB, 5 * MiB + 17, 3 * MiB, 1024, 8 * MiB - 1, 2 * MiB + MiB // 2, 0, 7 * MiB + 3]
    levels = np.array([1, 9, 3, 5, 2, 7, 4, 8], np.int32)
    datas = [dg.mix(s, 900 + i) if s else np.zeros(0, np.uint8) for i, s in enumerate(sizes)]
    offs = np.zeros(len(sizes), np.uint64)
    pos = 0
    for i, s in enumerate(sizes):
        offs[i] = pos
        pos += s + (-s) % 256
    pool = np.zeros(pos + 256, np.uint8)
    for i, d in enumerate(datas):
        pool[int(offs[i]):int(offs[i]) + d.size] = d
    t = torch.from_numpy(pool).to(dev)
    bz.set_max_batch_bytes(8 * MiB)  # sub-batches: 8M | 5M | 3M+1K | 8M-1 | 2.5M+0 | 7M
    try:
        _case(f"batch of {len(sizes)} inputs up to 8 MiB, 8 MiB sub-batches")
        ptr, bo, by = bz.compress_batch_device(t, offs, sizes, levels, 0)
        allout = _d2h(ptr, int(bo[-1] + by[-1]))
    finally:
        bz.set_max_batch_bytes(1 << 30)
    for i, d in enumerate(datas):
        got = allout[int(bo[i]):int(bo[i] + by[i])]
        lv = int(levels[i])
        _same(got, oracle.compress(d, lv, block_bytes=100000 * lv, threads=8), f"input {i} ({sizes[i]} B, L{lv})")


@pytest.mark.parametrize("overlap", [True, False])
def test_sub_batch_overlap_identical(dev, overlap):
    """Sub-batches on two overlapped pipelines (and on one) give the oracle's
    streams, with empty, tiny and multi-block inputs across 2 MiB
    sub-batches."""
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    sizes = [1500000, 0, 700000, 2 * MiB + 5, 1, 900001, 300000, 1200000, 64, 2 * MiB - 7]
    levels = np.array([9, 1, 2, 9, 5, 3, 1, 7, 9, 4], np.int32)
    datas = [dg.mix(s, 1300 + i) if s else np.zeros(0, np.uint8) for i, s in enumerate(sizes)]
    offs = np.cumsum([0] + [s + (-s) % 256 for s in sizes[:-1]]).astype(np.uint64)
    pool = np.zeros(int(offs[-1]) + sizes[-1] + 256, np.uint8)
    for i, d in enumerate(datas):
        pool[int(offs[i]):int(offs[i]) + d.size] = d
    t = torch.from_numpy(pool).to(dev)
    bz.set_max_batch_bytes(2 * MiB)
    bz.set_batch_overlap(overlap)
    try:
        _case(f"sub-batch overlap {overlap}: {len(sizes)} inputs, 2 MiB sub-batches")
        for rep in range(2):
            ptr, bo, by = bz.compress_batch_device(t, offs, sizes, levels, 0)
            allout = _d2h(ptr, int(bo[-1] + by[-1]))
            for i, d in enumerate(datas):
                lv = int(levels[i])
                _same(allout[int(bo[i]):int(bo[i] + by[i])],
                      oracle.compress(d, lv, block_bytes=100000 * lv, threads=8), f"rep {rep} input {i}")
    finally:
        bz.set_max_batch_bytes(1 << 30)
        bz.set_batch_overlap(True)


def _requests(n, seed, median):
    rng = np.random.default_rng(seed)
    sizes = np.minimum(16 * MiB, np.ceil(np.exp(rng.normal(np.log(median), 1.5, n)))).astype(np.int64)
    levels = rng.integers(1, 10, n).astype(np.int32)
    return sizes, levels, rng


def test_serving_64k_median(dev):
    """Requests of the serving workload's size law (log-normal, median
    64 KiB, sigma 1.5, capped at 16 MiB) under bursty arrivals: every
    result equals the oracle's."""
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    n = 96
    sizes, levels, rng = _requests(n, 64, 64 * 1024)
    pool = dg.mix(64 * MiB, 640)
    offs = (rng.random(n) * (pool.size - sizes)).astype(np.int64)
    d_pool = torch.from_numpy(pool).to(dev)
    arena = torch.zeros(int(sizes.sum() * 2 + n * 64 + MiB), dtype=torch.uint8, device=dev)
    srv = bz.Server(mode="batched", workers=4, max_batch_bytes=32 * MiB)
    srv.keep_outputs(arena)
    arrival = np.cumsum(rng.exponential(50000, n)).astype(np.int64)
    _case(f"serve 64 KiB median n={n} total={int(sizes.sum())}")
    done, ob, oo, bs = srv.replay(d_pool, offs, sizes, levels, arrival)
    srv.close()
    assert (done >= arrival).all()
    assert bs.max() > 1, "no request shared a pipeline run"
    host = arena.cpu().numpy().tobytes()
    for i in range(n):
        got = host[int(oo[i]):int(oo[i] + ob[i])]
        src = pool[int(offs[i]):int(offs[i] + sizes[i])]
        lv = int(levels[i])
        _same(got, oracle.compress(src, lv, block_bytes=100000 * lv, threads=4), f"request {i}")


def test_live_requests_then_replay(dev):
    """Live requests still running when a replay starts: the replay waits
    for them, their tickets stay valid, and both sets of results are right;
    a ticket waited for twice is refused."""
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    live = [dg.mix(s, 700 + s) for s in (3 * MiB, 200000, 5000)]
    bufs = [torch.from_numpy(d).to(dev) for d in live]
    sizes, levels, rng = _requests(24, 65, 64 * 1024)
    pool = dg.mix(16 * MiB, 650)
    offs = (rng.random(sizes.size) * (pool.size - sizes)).astype(np.int64)
    d_pool = torch.from_numpy(pool).to(dev)
    arena = torch.zeros(int(sizes.sum() * 2 + 16 * MiB), dtype=torch.uint8, device=dev)
    srv = bz.Server(mode="batched", workers=2, max_batch_bytes=8 * MiB)
    srv.keep_outputs(arena)
    _case("serve: live requests, then a replay")
    tickets = [srv.submit(b, 9) for b in bufs]
    arrival = np.cumsum(rng.exponential(20000, sizes.size)).astype(np.int64)
    done, ob, oo, bs = srv.replay(d_pool, offs, sizes, levels, arrival)
    live_res = [srv.wait(t) for t in tickets]
    with pytest.raises(bz.BzgError):
        srv.wait(tickets[0])
    srv.close()
    host = arena.cpu().numpy().tobytes()
    for d, (nb, off, lat) in zip(live, live_res):
        assert off >= 0
        _same(host[off:off + nb], oracle.compress(d, 9, block_bytes=900000), f"live request {d.size} B")
    for i in range(sizes.size):
        src = pool[int(offs[i]):int(offs[i] + sizes[i])]
        lv = int(levels[i])
        _same(host[int(oo[i]):int(oo[i] + ob[i])], oracle.compress(src, lv, block_bytes=100000 * lv),
              f"replayed request {i}")
