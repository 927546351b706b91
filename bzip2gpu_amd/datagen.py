"""Seeded synthetic inputs of the benchmark workloads (wraps libbzgdata.so).

Workloads (seed = 20261203 + workload index, sub-seeded per segment/input):
  text1   1 MiB order-1 Markov text, Zipf 1.1 rows (one input, level 9)
  mix     256 MiB of 1 MiB segments: 50% order-2 text (Zipf 1.2),
          25% uniform bytes, 25% runs with geometric lengths (mean 64)
  batch   4096 inputs, sizes log-uniform in 1 KiB..8 MiB, levels 1..9, mix content
  periodic 64 MiB of 900000-byte blocks repeating a seeded string of period
          p in (1, 2, 7, 1000, 65536) by block, 0.1% single-byte mutations
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libbzgdata.so")
_lib = None
SEED0 = 20261203


def _L():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            raise RuntimeError(f"{_SO} missing; run bzip2gpu_amd.build()")
        L = ctypes.CDLL(_SO)
        P, sz, u64, d, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_double, ctypes.c_int
        L.dg_markov.argtypes = [P, sz, u64, i, d]
        L.dg_uniform.argtypes = [P, sz, u64]
        L.dg_runs.argtypes = [P, sz, u64, d]
        L.dg_mix.argtypes = [P, sz, u64]
        L.dg_periodic.argtypes = [P, sz, u64, sz, d]
        for f in ("dg_markov", "dg_uniform", "dg_runs", "dg_mix", "dg_periodic"):
            getattr(L, f).restype = None
        _lib = L
    return _lib


def _buf(n):
    a = np.empty(n, dtype=np.uint8)
    return a, a.ctypes.data_as(ctypes.c_void_p)


def markov(n: int, seed: int, order: int = 1, s: float = 1.1) -> np.ndarray:
    a, p = _buf(n)
    _L().dg_markov(p, n, seed, order, s)
    return a


def uniform(n: int, seed: int) -> np.ndarray:
    a, p = _buf(n)
    _L().dg_uniform(p, n, seed)
    return a


def runs(n: int, seed: int, mean: float = 64.0) -> np.ndarray:
    a, p = _buf(n)
    _L().dg_runs(p, n, seed, mean)
    return a


def mix(n: int, seed: int) -> np.ndarray:
    a, p = _buf(n)
    _L().dg_mix(p, n, seed)
    return a


def periodic(n: int, seed: int, period: int, mut_frac: float = 0.001) -> np.ndarray:
    a, p = _buf(n)
    _L().dg_periodic(p, n, seed, period, mut_frac)
    return a


def workload_text1(n: int = 1 << 20) -> np.ndarray:
    return markov(n, SEED0 + 0, order=1, s=1.1)


def workload_mix(n: int = 256 << 20) -> np.ndarray:
    return mix(n, SEED0 + 1)


def workload_batch(n_inputs: int = 4096, max_bytes: int = 8 << 20, seed: int = SEED0 + 2):
    """Returns (list of arrays, levels)."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(1024), np.log(max_bytes), n_inputs))).astype(np.int64)
    levels = rng.integers(1, 10, n_inputs).astype(np.int32)
    data = [mix(int(sz), seed * 7 + i) for i, sz in enumerate(sizes)]
    return data, levels


def workload_periodic(n: int = 64 << 20, block: int = 900000, seed: int = SEED0 + 3) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    periods = (1, 2, 7, 1000, 65536)
    for j, off in enumerate(range(0, n, block)):
        ln = min(block, n - off)
        out[off:off + ln] = periodic(ln, seed * 131 + j, periods[j % 5])
    return out
