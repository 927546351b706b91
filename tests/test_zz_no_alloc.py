"""Allocation-free batched entry point (bzg_compress_batch_device_noalloc):
its streams equal those of bzg_compress_batch_device on the same inputs, on
consecutive calls (the second smaller, on the array the first left behind),
at the limits it documents, and batches beyond them are refused."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _pool(sizes, seed):
    import bzip2gpu_amd.datagen as dg
    offs, pos = [], 0
    for s in sizes:
        offs.append(pos)
        pos += s + (-s) % 256
    pool = np.zeros(pos + 256, np.uint8)
    for i, s in enumerate(sizes):
        if s:
            pool[offs[i]:offs[i] + s] = dg.mix(s, seed + i)
    return pool, np.array(offs, np.uint64)


def _run_both(dev, sizes, levels, seed, block_bytes=0):
    import torch
    import bzip2gpu_amd as bz
    pool, offs = _pool(sizes, seed)
    t = torch.from_numpy(pool).to(dev)
    # worst-case stream sizes: payload room plus about 6.5 KB of header words per block
    out = torch.empty(int(sum(sizes) * 2 + 8192 * (len(sizes) + 2) + MiB), dtype=torch.uint8, device=dev)
    lv = np.array(levels, np.int32)
    bo, by = bz.compress_batch_device_noalloc(t, offs, sizes, lv, out, block_bytes)
    got = out.cpu().numpy().tobytes()
    ptr, rbo, rby = bz.compress_batch_device(t, offs, sizes, lv, block_bytes)
    import ctypes
    torch.cuda.synchronize()
    end = int(rbo[-1] + rby[-1])
    ref = (ctypes.c_uint8 * max(1, end))()
    assert ctypes.CDLL("libamdhip64.so").hipMemcpy(ref, ctypes.c_void_p(ptr), ctypes.c_size_t(end), 2) == 0
    ref = bytes(ref)
    for i in range(len(sizes)):
        a = got[int(bo[i]):int(bo[i] + by[i])]
        b = ref[int(rbo[i]):int(rbo[i] + rby[i])]
        assert a == b, f"input {i} ({sizes[i]} B): allocation-free stream differs from the batched entry point's"
    return pool, offs, got, bo, by


def test_noalloc_matches_batched_entry_point(dev):
    # first call: near the 1 MiB limit; second call: smaller, on the array
    # the first call left behind
    _case("no-alloc call 1: 3 inputs, 1 MiB")
    sizes1 = [600000, 300000, MiB - 900000]
    pool, offs, got, bo, by = _run_both(dev, sizes1, [9, 3, 1], 10)
    lv = [9, 3, 1]
    for i, s in enumerate(sizes1):  # and the oracle, for one input
        if i == 1:
            want = oracle.compress(pool[int(offs[i]):int(offs[i]) + s], lv[i], block_bytes=100000 * lv[i])
            assert got[int(bo[i]):int(bo[i] + by[i])] == want
    _case("no-alloc call 2: 2 inputs, smaller")
    _run_both(dev, [200000, 1], [5, 9], 20)


def test_noalloc_at_documented_limits(dev):
    # 64 blocks: 62 one-byte inputs and one input cut into two blocks
    _case("no-alloc: 64 blocks, 1 MiB")
    sizes = [1] * 62 + [MiB - 62]
    _run_both(dev, sizes, [9] * 63, 30, block_bytes=900000)
    _case("no-alloc: 64 inputs of 16 KiB at level 1")
    _run_both(dev, [16384] * 64, [1] * 64, 40)


def test_noalloc_refuses_beyond_limits(dev):
    import torch
    import bzip2gpu_amd as bz
    t = torch.zeros(2 * MiB, dtype=torch.uint8, device=dev)
    out = torch.empty(4 * MiB, dtype=torch.uint8, device=dev)
    _case("no-alloc: too many bytes / blocks / too small output")
    with pytest.raises(bz.BzgError, match="at most"):
        bz.compress_batch_device_noalloc(t, [0], [MiB + 1], [9], out)
    with pytest.raises(bz.BzgError, match="at most"):
        bz.compress_batch_device_noalloc(t, np.zeros(65, np.uint64), [1] * 65, [9] * 65, out)
    with pytest.raises(bz.BzgError, match="too small"):
        bz.compress_batch_device_noalloc(t, [0], [MiB], [9], out[:1000])
    # a valid call after the refusals
    _run_both(dev, [123457], [2], 50)
