"""Serving loop: every request's stream is the same whether it was served
alone or batched with others, and equals the oracle's."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _d2h(t, off, n):
    return t[off:off + n].cpu().numpy().tobytes()


def _requests(n, seed):
    rng = np.random.default_rng(seed)
    sizes = np.minimum(16 << 20, np.ceil(np.exp(rng.normal(np.log(16384), 1.5, n)))).astype(np.int64)
    sizes[0] = 0  # an empty request in the mix
    levels = rng.integers(1, 10, n).astype(np.int32)
    return sizes, levels, rng


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_replay_alone_equals_mixed(dev, mode):
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    n = 120 if mode == "batched" else 24
    sizes, levels, rng = _requests(n, 7)
    pool = dg.mix(32 << 20, 77)
    offs = (rng.random(n) * (pool.size - sizes)).astype(np.int64)
    d_pool = torch.from_numpy(pool).to(dev)
    arena = torch.zeros(int(sizes.sum() * 2 + n * 64 + (1 << 20)), dtype=torch.uint8, device=dev)
    srv = bz.Server(mode=mode, workers=3, max_batch_bytes=8 << 20)
    srv.keep_outputs(arena)
    # bursty arrivals so that several requests share a pipeline run
    arrival = np.cumsum(rng.exponential(20000, n)).astype(np.int64)
    _case(f"serve replay mode={mode} n={n}")
    done, ob, oo, bs = srv.replay(d_pool, offs, sizes, levels, arrival)
    srv.close()
    assert (done >= arrival).all()
    if mode == "batched":
        assert bs.max() > 1, "no request shared a pipeline run"
    for i in range(n):
        got = _d2h(arena, int(oo[i]), int(ob[i]))
        src = pool[int(offs[i]):int(offs[i] + sizes[i])]
        alone = bz.compress(src.tobytes(), int(levels[i]), block_bytes=100000 * int(levels[i]))
        assert got == alone, f"request {i} differs from its alone result"
        if i % 10 == 0:
            assert got == oracle.compress(src, int(levels[i]), block_bytes=100000 * int(levels[i])), i


def test_submit_wait_live(dev):
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    data = [dg.mix(s, 300 + s) for s in (1000, 250000, 1 << 20)]
    bufs = [torch.from_numpy(d).to(dev) for d in data]
    arena = torch.zeros(8 << 20, dtype=torch.uint8, device=dev)
    srv = bz.Server(mode="batched", workers=2)
    srv.keep_outputs(arena)
    _case("serve live submit/wait")
    tickets = [srv.submit(b, 9) for b in bufs]
    for d, t in zip(data, tickets):
        nb, off, lat = srv.wait(t)
        assert lat > 0
        assert _d2h(arena, off, nb) == oracle.compress(d, 9, block_bytes=900000)
    srv.close()
