#!/usr/bin/env python3
"""Benchmark: bzip2 compression input GB/s on one MI355X.

Default workload (batch4096): 4096 independent inputs already resident in
HBM, sizes floor(exp(U(ln 1 KiB, ln 8 MiB))) (about 3.84 GB in all), levels
U{1..9}, each cut into 100000*level-byte blocks, content = slices of a
seeded 256 MiB pool of 1 MiB segments drawn 50/25/25 from order-2 Markov
text (Zipf 1.2), uniform bytes and runs with geometric lengths of mean 64.
A step compresses every input with the batched pipeline (all blocks of a
sub-batch of at most 1280 MiB in one device pipeline, streams assembled on the
device).  Beside it, on the same inputs: the plain per-block path on one
stream, the plain path with the reference's own concurrency (16 host
threads, one HIP stream each, one block per pipeline run), and the CPU
oracle and stock libbz2 on the host cores; every GPU output is checked
byte-for-byte against the others and against the oracle.

Other workloads: mix256 (one 256 MiB stream, level 9, 900000-byte blocks),
text1, periodic64, serve, decompress.

Prints one JSON line (see the keys below).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MiB = 1 << 20


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="batch4096",
                   choices=["mix256", "text1", "periodic64", "batch4096", "serve", "decompress",
                            "decompress-batch"])
    p.add_argument("--inputs", type=int, default=4096, help="batch4096: number of inputs")
    p.add_argument("--lambdas", default="2000,8000,32000", help="serve: Poisson arrival rates (req/s)")
    p.add_argument("--serve-seconds", type=float, default=60.0, help="serve: arrival window per rate")
    p.add_argument("--plain-seconds", type=float, default=1.0, help="serve: arrival window of the plain-path run")
    p.add_argument("--workers", type=int, default=4, help="serve: pipeline workers (HIP streams)")
    p.add_argument("--mib", type=int, default=0, help="override input size (MiB)")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--max-batch-mib", type=int, default=1280, help="batch: raw MiB per device sub-batch")
    p.add_argument("--no-overlap", action="store_true", help="batch: run sub-batches one after another")
    p.add_argument("--tune-force", default="",
                   help="pin autotuner variants instead of timing them, e.g. 0=0,1=0 (knob=variant)")
    p.add_argument("--plain-threads", type=int, default=16,
                   help="batch4096: host threads of the reference-style plain path (0: skip)")
    p.add_argument("--cpu-sample-mib", type=int, default=256,
                   help="batch4096: raw MiB of inputs (taken in order) the CPU baselines compress per run")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--probe", default="k_lscatter", help="kernel timed live with HIP events for the roofline")
    p.add_argument("--calib-copy", action="store_true",
                   help="also run one plain device copy of the input (PMC traffic calibration)")
    p.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                   help="JSON with HBM bytes per launch of the probed kernel, measured by rocprofv3 PMC runs")
    return p.parse_args()


# Algorithmic HBM bytes of one unit of work of each probed kernel, and the
# name of the unit.
ALG_BYTES = {
    # one LSD radix pass of the large-group suffix sort that reads keys: read
    # key 4 + SA 4, write key 4 + SA 4 per row and pass (round 0's first pass
    # reads the text instead and is not probed)
    "k_lscatter": (16, "radix_pass_rows"),
    # LDS sort of a pack: read key 4 + SA 4, write SA 4 + scatter rank 4
    "k_sort_small": (16, "bwt_rows_small"),
    # MTF: read last-column byte, write MTF byte
    "k_mtf_apply": (2, "rle1_bytes"),
}


def _workload(name, mib):
    import bzip2gpu_amd.datagen as dg
    if name == "mix256":
        n = (mib or 256) * MiB
        return dg.mix(n, dg.SEED0 + 1), {"workload": "mix", "bytes": n, "level": 9, "block_bytes": 900000,
                                         "content": "1MiB segments: 50% order-2 text Zipf1.2 / 25% uniform / 25% runs mean 64"}
    if name == "text1":
        n = (mib or 1) * MiB
        return dg.markov(n, dg.SEED0, order=1, s=1.1), {"workload": "text1", "bytes": n, "level": 9,
                                                        "block_bytes": 900000, "content": "order-1 Markov text Zipf1.1"}
    n = (mib or 64) * MiB
    return dg.workload_periodic(n), {"workload": "periodic", "bytes": n, "level": 9, "block_bytes": 900000,
                                     "content": "period 1/2/7/1000/65536 by block, 0.1% mutations"}


def _batch_inputs(n_inputs, pool_bytes, seed=20261205):
    """batch4096: input i is a slice of the seeded mix pool; size
    floor(exp(U(ln 1024, ln 8 MiB))), level U{1..9}, blocks of 100000*level."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(1024), np.log(8 * MiB), n_inputs))).astype(np.int64)
    levels = rng.integers(1, 10, n_inputs).astype(np.int32)
    offs = (rng.random(n_inputs) * (pool_bytes - sizes)).astype(np.int64)
    return offs, sizes, levels


def _cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def _copy_peak_gbps():
    """Achievable HBM bandwidth on this box: a 1 GiB device-to-device copy
    (read + write), best of 5."""
    import torch
    a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    best = None
    for _ in range(5):
        t = time.perf_counter()
        b.copy_(a)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    del a, b
    return 2 * (1 << 30) / best / 1e9


def _alg_bytes(tot):
    """Algorithmic HBM bytes of a batch: per block
    2n + 9m + 20 m R + 12 s + W/8 (CRC and RLE1, suffix sort with R
    prefix-doubling rounds, last column and MTF, RLE2, Huffman selection and
    encoding, payload), plus W/4 for the stream assembly on the device."""
    return (2 * tot["n"] + 9 * tot["m"] + 20 * tot["m_rounds"] + 12 * tot["s"] + tot["bits"] / 8
            + tot["bits"] / 4)


def _median_runs(fn, runs=5):
    fn()  # warm-up
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    ts.sort()
    return ts[len(ts) // 2], ts


def _cpu_baselines_batch(pool, offs, sizes, levels, sample_bytes):
    """The oracle (the reference path restated in C) and stock libbz2 (Python's
    bz2, libbz2 1.0.8), each over the first inputs of the batch up to
    sample_bytes, input-parallel on min(nproc, 16) host threads (16 is the
    reference's worker cap); 1 warm-up + median of 5 runs."""
    import bz2
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    threads = max(1, min(os.cpu_count() or 1, 16))
    k, acc = 0, 0
    while k < len(sizes) and acc + int(sizes[k]) <= sample_bytes:
        acc += int(sizes[k])
        k += 1
    k = max(k, 1)
    acc = int(sizes[:k].sum())
    items = [(pool[int(offs[i]):int(offs[i] + sizes[i])], int(levels[i])) for i in range(k)]
    outs = {}

    def run_oracle():
        with ThreadPoolExecutor(threads) as ex:
            res = list(ex.map(lambda it: oracle.compress(it[0], it[1], block_bytes=100000 * it[1]), items))
        outs["oracle"] = res

    def run_bz2():
        with ThreadPoolExecutor(threads) as ex:
            res = list(ex.map(lambda it: bz2.compress(it[0].tobytes(), it[1]), items))
        outs["bz2"] = res

    mo, ro = _median_runs(run_oracle)
    mb, rb = _median_runs(run_bz2)
    sample = (f"the first {k} of the batch's inputs ({acc} raw bytes, levels and 100000*level blocks as in "
              f"the batch), input-parallel on {threads} host threads; 1 warm-up + median of 5 runs")
    base = {"value": round(acc / mo / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": "oracle: " + sample, "runs_s": [round(x, 3) for x in ro]}
    base.update(_cpu_info())
    base["libbz2"] = {"value": round(acc / mb / 1e9, 4), "unit": "GB/s", "cores": threads,
                      "sample": "Python bz2 (libbz2 1.0.8, BZ2_bzBuffToBuffCompress semantics): " + sample,
                      "runs_s": [round(x, 3) for x in rb]}
    return base, outs["oracle"], outs["bz2"], k


def run_batch(a):
    import numpy as np
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    pool = dg.mix(256 * MiB, dg.SEED0 + 2)
    offs, sizes, levels = _batch_inputs(a.inputs, pool.size)
    offs_u, sizes_u = offs.astype(np.uint64), sizes.astype(np.uint64)
    total = int(sizes.sum())
    d_pool = torch.from_numpy(pool).cuda()
    torch.cuda.synchronize()
    bz.set_max_batch_bytes(a.max_batch_mib * MiB)
    bz.set_batch_overlap(not a.no_overlap)

    def batched():
        return bz.compress_batch_device(d_pool, offs_u, sizes_u, levels, 0)

    tune = _autotune(batched, forced=a.tune_force)
    for _ in range(max(0, a.warmup)):
        batched()
    torch.cuda.synchronize()
    bz.probe_kernel(a.probe)
    bz.probe_read()
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        ptr, bo, by = batched()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    probe_ms, probe_n = bz.probe_read()
    bz.probe_kernel(None)
    stages, rounds, rows = bz.last_stage_times()
    totals = bz.last_batch_totals()
    small_rows, large_rows, pass_rows = bz.last_sort_rows()
    step_s = wall / a.steps
    value = total / step_s / 1e9
    end = int(bo[-1] + by[-1])
    allout = _hip_copy_d2h(ptr, end)
    outs = [allout[int(bo[i]):int(bo[i] + by[i])] for i in range(a.inputs)]
    res = {
        "metric": "input GB/s at 1 MI355X, batched vs plain per-block (bit-exact); p50/p99 latency",
        "value": round(value, 4), "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(step_s * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "batch", "inputs": a.inputs, "bytes": total, "blocks": totals["blocks"],
                   "sizes": "floor(exp(U(ln 1 KiB, ln 8 MiB)))", "levels": "U{1..9}", "block_bytes": "100000*level",
                   "content": "slices of a seeded 256 MiB pool of 1 MiB segments: 50% order-2 text Zipf1.2 / "
                              "25% uniform / 25% runs mean 64",
                   "path": f"batched (sub-batches of <= {a.max_batch_mib} MiB raw"
                           + (", one after another)" if a.no_overlap else
                              ", alternating between two pipelines on two HIP streams, sorts in order)")},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "stage_ms": {k: round(v, 3) for k, v in stages.items()}, "bwt_rounds": rounds,
        "compressed_bytes": int(sum(by)), "autotune": tune,
        "sort_rows": {"small_groups": small_rows, "large_groups": large_rows, "large_x_passes": pass_rows},
    }
    per_unit, unit = ALG_BYTES.get(a.probe, (0, ""))
    units = {"radix_pass_rows": pass_rows}.get(unit, 0)
    copy_peak = _copy_peak_gbps()
    if probe_n and units and per_unit:
        per_launch = per_unit * units * a.steps / probe_n
        achieved = per_launch / (probe_ms / probe_n / 1e3) / 1e9
        traffic = None
        if a.traffic and os.path.exists(a.traffic):
            with open(a.traffic) as f:
                traffic = json.load(f).get(f"batch:{a.probe}")
        res["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                           "frac": round(achieved / 8000.0, 4), "traffic": traffic,
                           "kernel": a.probe, "launches": probe_n,
                           "avg_launch_us": round(probe_ms / probe_n * 1e3, 2),
                           "alg_bytes_per_launch": int(per_launch), "alg_bytes_rule": f"{per_unit} B x {unit}"}
    alg = _alg_bytes(totals)
    pipe = alg / step_s / 1e9
    res.setdefault("roofline", {})["pipeline"] = {
        "alg_bytes_per_step": int(alg), "achieved": round(pipe, 2), "frac": round(pipe / 8000.0, 4),
        "copy_peak_gbps": round(copy_peak, 1), "frac_of_copy_peak": round(pipe / copy_peak, 4),
        "rule": "per block 2n + 9m + 20 m R + 12 s + W/8, + W/4 assembly", "totals": totals}
    if not a.no_plain:
        torch.cuda.synchronize()
        t = time.perf_counter()
        same = True
        for i in range(a.inputs):
            o = bz.compress_plain_device(d_pool[int(offs[i]):int(offs[i] + sizes[i])], int(levels[i]),
                                         100000 * int(levels[i]))
            same = same and (o == outs[i])
        dtp = time.perf_counter() - t
        res["plain"] = {"value": round(total / dtp / 1e9, 4), "unit": "GB/s", "seconds": round(dtp, 3),
                        "path": "one block per pipeline run, one HIP stream, host serialisation",
                        "identical_to_batched": bool(same)}
        res["batched_over_plain"] = round(value / (total / dtp / 1e9), 2)
    if a.plain_threads > 0:
        bz.compress_plain_threads(d_pool, offs_u[:64], sizes_u[:64], levels[:64], 0, a.plain_threads)  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        pt = bz.compress_plain_threads(d_pool, offs_u, sizes_u, levels, 0, a.plain_threads)
        dtt = time.perf_counter() - t
        res["plain_threads"] = {"value": round(total / dtt / 1e9, 4), "unit": "GB/s", "seconds": round(dtt, 3),
                                "threads": a.plain_threads,
                                "path": "the reference's concurrency: host threads with one HIP stream each, one "
                                        "block per pipeline run, host serialisation",
                                "identical_to_batched": pt == outs}
        res["batched_over_plain_threads"] = round(value / (total / dtt / 1e9), 2)
    if not a.no_cpu:
        cpu, o_outs, b_outs, k = _cpu_baselines_batch(pool, offs, sizes, levels, a.cpu_sample_mib * MiB)
        res["cpu_baseline"] = cpu
        res["identical_to_oracle"] = all(o_outs[i] == outs[i] for i in range(k))
        res["oracle_checked_inputs"] = k
        res["libbz2_identical_inputs"] = sum(1 for i in range(k) if b_outs[i] == outs[i])
    elif not a.no_check:
        import oracle
        idx = np.random.default_rng(5).choice(a.inputs, size=min(64, a.inputs), replace=False)
        res["identical_to_oracle"] = all(
            oracle.compress(pool[int(offs[i]):int(offs[i] + sizes[i])], int(levels[i]),
                            block_bytes=100000 * int(levels[i]), threads=8) == outs[int(i)] for i in idx)
        res["oracle_checked_inputs"] = int(idx.size)
    print(json.dumps(res), flush=True)


def _serve_requests(lam, seconds, pool_bytes, seed):
    """Poisson arrivals at lam req/s for `seconds`; size min(16 MiB,
    ceil(exp(N(ln 64 KiB, 1.5)))), level U{1..9}, content = a slice of the
    seeded mix pool."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n = max(1, int(rng.poisson(lam * seconds)))
    arrival = np.sort(rng.random(n) * seconds * 1e9).astype(np.int64)
    sizes = np.minimum(16 * MiB, np.ceil(np.exp(rng.normal(np.log(64 * 1024), 1.5, n)))).astype(np.int64)
    levels = rng.integers(1, 10, n).astype(np.int32)
    offs = (rng.random(n) * (pool_bytes - sizes)).astype(np.int64)
    return arrival, sizes, levels, offs


def _serve_stats(arrival, done, sizes):
    import numpy as np
    lat = (done - arrival) / 1e6
    span = (done.max() - arrival.min()) / 1e9
    return {"requests": int(arrival.size), "bytes": int(sizes.sum()),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "gbps": round(float(sizes.sum()) / span / 1e9, 4),
            "offered_gbps": round(float(sizes.sum()) / (arrival.max() / 1e9) / 1e9, 4) if arrival.max() > 0 else None}


def run_serve(a):
    import numpy as np
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg
    pool = dg.mix(256 * MiB, dg.SEED0 + 4)
    d_pool = torch.from_numpy(pool).cuda()
    torch.cuda.synchronize()
    lams = [int(x) for x in a.lambdas.split(",") if x]
    srv = bz.Server(mode="batched", workers=a.workers, max_batch_bytes=256 * MiB, hbm_budget_bytes=200 * 10**9)
    # warm-up: grow every worker's workspace
    arr, sz, lv, of = _serve_requests(lams[-1], 0.5, pool.size, 1)
    for _ in range(max(0, a.warmup)):
        srv.replay(d_pool, of, sz, lv, arr)
    per = []
    for i, lam in enumerate(lams):
        arr, sz, lv, of = _serve_requests(lam, a.serve_seconds, pool.size, 100 + i)
        done, ob, oo, bs = srv.replay(d_pool, of, sz, lv, arr)
        st = _serve_stats(arr, done, sz)
        st.update({"lambda": lam, "mean_batch": round(float(bs.mean()), 2), "max_batch": int(bs.max())})
        per.append(st)
        print(f"[serve] lambda {lam}/s done: {st}", file=sys.stderr, flush=True)
    srv.close()
    # plain path at the lowest rate, shorter window (it cannot keep up)
    psrv = bz.Server(mode="plain", workers=1)
    arr, sz, lv, of = _serve_requests(lams[0], a.plain_seconds, pool.size, 999)
    done, ob, oo, bs = psrv.replay(d_pool, of, sz, lv, arr)
    psrv.close()
    plain = _serve_stats(arr, done, sz)
    plain["lambda"] = lams[0]
    plain["window_s"] = a.plain_seconds
    # the batched server on the same short plain-path trace (equal load)
    srv = bz.Server(mode="batched", workers=a.workers, max_batch_bytes=256 * MiB, hbm_budget_bytes=200 * 10**9)
    done2, _, _, _ = srv.replay(d_pool, of, sz, lv, arr)
    srv.close()
    same_load = _serve_stats(arr, done2, sz)
    top = per[-1]
    res = {
        "metric": "input GB/s at 1 MI355X, batched vs plain per-block (bit-exact); p50/p99 latency",
        "value": top["gbps"], "unit": "GB/s", "n_gpus": 1, "steps": len(lams), "warmup": a.warmup,
        "ms_per_step": round(a.serve_seconds * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "serve", "lambdas": lams, "window_s": a.serve_seconds,
                   "sizes": "min(16 MiB, lognormal(median 64 KiB, sigma 1.5))", "levels": "U{1..9}",
                   "block_bytes": "100000*level", "hbm_budget_bytes": 200 * 10**9, "workers": a.workers,
                   "max_batch_bytes": 256 * MiB, "path": "continuous batching"},
        "latency_ms": {"p50": top["p50_ms"], "p99": top["p99_ms"]},
        "per_lambda": per,
        "plain_equal_load": plain,
        "batched_equal_load": same_load,
    }
    print(json.dumps(res), flush=True)


def _autotune(run_once, max_runs=16, forced=""):
    """Untimed set-up: lets the library's autotuner time each launch shape of
    the radix scatter (knob 0), then each launch policy of the suffix-sort
    rounds (knob 1), and each MTF segment run length (knob 2), on this
    workload until every size class it met has
    settled (each variant is timed twice per class).  `forced` ("knob=variant,
    ...") pins variants instead (profiling runs)."""
    import torch
    import bzip2gpu_amd as bz

    if forced:
        pins = dict(tuple(int(x) for x in kv.split("=")) for kv in forced.split(","))
        for k, v in pins.items():
            bz.tune_force(k, v)
        return {"forced": {f"knob{k}": v for k, v in pins.items()}}

    def seen(knob):
        out = [(c, bz.tune_state(knob, c)) for c in range(4)]
        return [(c, st) for c, st in out if any(n for _, _, n in st[1]) or st[0] >= 0]

    def unsettled():
        return any(not seen(k) or any(st[0] < 0 for _, st in seen(k)) for k in (0, 1, 2))

    runs = 0
    while runs < max_runs and unsettled():
        run_once()
        runs += 1
    torch.cuda.synchronize()
    out = {"setup_runs": runs}
    names = {0: ("radix_scatter_threads", "ns_per_row_pass"), 1: ("sort_launch", "ns_per_sorted_row"),
             2: ("mtf_segments_per_wave", "ns_per_capacity_tile")}
    for k in (0, 1, 2):
        for c, (chosen, rows) in seen(k):
            key, unit = names[k]
            val = rows[chosen][0] if chosen >= 0 else None
            if k == 1 and val is not None:
                val = ("per_round", "device_driven_x1", "device_driven_x4")[val]
            out[f"knob{k}_class{c}"] = {key: val, unit: {str(v): round(ns, 4) for v, ns, _ in rows}}
    return out


def run_decompress(a):
    """Decompression of the mix workload's .bz2 stream (written once by the
    batched compressor, untimed) with the compressed bytes resident in HBM;
    value = raw output GB/s.  The output is checked against the input, and
    stock libbz2 (Python's bz2, one thread) is timed on the same stream."""
    import bz2
    import numpy as np
    import torch
    import bzip2gpu_amd as bz

    data, cfg = _workload("mix256", a.mib)
    n = data.size
    comp = bz.compress(data.tobytes(), 9, cfg["block_bytes"])
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(np.frombuffer(comp, dtype=np.uint8).copy()).to(dev)
    torch.cuda.synchronize()
    for _ in range(max(0, a.warmup)):
        bz.decompress_device(d_in)
    torch.cuda.synchronize()
    if a.probe:
        bz.probe_kernel(a.probe)
        bz.probe_read()
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        ptr, nout = bz.decompress_device(d_in)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    wall = time.perf_counter() - t0
    probe_ms, probe_n = bz.probe_read() if a.probe else (0.0, 0)
    bz.probe_kernel(None)
    stages, nblk, nstr = bz.last_decode_times()
    value = n / (wall / a.steps) / 1e9
    res = {
        "metric": "decompression output GB/s at 1 MI355X (bit-exact, CRC-checked)",
        "value": round(value, 4), "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": dict(cfg, workload="decompress-mix", compressed_bytes=len(comp), blocks=nblk),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "stage_ms": dict(zip(["huffman", "mtf_rle2", "fl_sort", "bwt_walk", "rle1_crc"],
                             [round(x, 3) for x in stages])),
    }
    if probe_n:
        res["probe"] = {"kernel": a.probe, "launches": probe_n, "avg_launch_us": round(probe_ms / probe_n * 1e3, 2)}
    if not a.no_check:
        res["identical_to_input"] = nout == n and _hip_copy_d2h(ptr, nout) == data.tobytes()
    if not a.no_cpu:
        # stock libbz2 refuses some blocks of this stream (RLE1 beyond
        # 900000 bytes under the fixed 900000-byte cut), so it decodes its
        # own level-9 stream of a 32 MiB sample of the same input
        sample = data[: min(n, 32 * MiB)].tobytes()
        stock = bz2.compress(sample, 9)
        t = time.perf_counter()
        ref = bz2.decompress(stock)
        dt = time.perf_counter() - t
        res["cpu_baseline"] = {"value": round(len(sample) / dt / 1e9, 4), "unit": "GB/s", "cores": 1,
                               "kind": "libbz2",
                               "sample": f"Python bz2 (libbz2 1.0.8) decoding the stock level-9 stream of the first "
                                         f"{len(sample) // MiB} MiB of the input, {dt:.2f} s wall",
                               "identical": ref == sample}
    print(json.dumps(res), flush=True)


def run_decompress_batch(a):
    """Decompression of the batched workload: its inputs compressed once by
    the batched path (untimed), the streams concatenated into one buffer
    resident in HBM, all blocks of all streams decoded together.  value =
    raw output GB/s; the output is checked against the inputs."""
    import bz2
    import numpy as np
    import torch
    import bzip2gpu_amd as bz
    import bzip2gpu_amd.datagen as dg

    pool = dg.mix(256 * MiB, dg.SEED0 + 2)
    offs, sizes, levels = _batch_inputs(a.inputs, pool.size)
    d_pool = torch.from_numpy(pool).cuda()
    ptr, bo, by = bz.compress_batch_device(d_pool, offs.astype(np.uint64), sizes.astype(np.uint64), levels, 0)
    torch.cuda.synchronize()
    allout = _hip_copy_d2h(ptr, int(bo[-1] + by[-1]))
    comp = b"".join(allout[int(bo[i]):int(bo[i] + by[i])] for i in range(a.inputs))
    del allout
    total = int(sizes.sum())
    d_in = torch.from_numpy(np.frombuffer(comp, dtype=np.uint8).copy()).cuda()
    torch.cuda.synchronize()
    for _ in range(max(0, a.warmup)):
        bz.decompress_device(d_in)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        optr, nout = bz.decompress_device(d_in)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    wall = time.perf_counter() - t0
    stages, nblk, nstr = bz.last_decode_times()
    res = {
        "metric": "decompression output GB/s at 1 MI355X (bit-exact, CRC-checked)",
        "value": round(total / (wall / a.steps) / 1e9, 4), "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "decompress-batch", "streams": nstr, "blocks": nblk, "bytes": total,
                   "compressed_bytes": len(comp),
                   "content": "the batched workload's 4096 inputs (levels 1-9, 100000*level blocks), compressed "
                              "by the batched path, streams concatenated"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "stage_ms": dict(zip(["huffman", "mtf_rle2", "fl_sort", "bwt_walk", "rle1_crc"],
                             [round(x, 3) for x in stages])),
    }
    if not a.no_check:
        got = _hip_copy_d2h(optr, nout)
        want = b"".join(pool[int(offs[i]):int(offs[i] + sizes[i])].tobytes() for i in range(a.inputs))
        res["identical_to_input"] = nout == total and got == want
    if not a.no_cpu:
        # stock libbz2 on one host core, decoding its own streams of the
        # first inputs (up to 32 MiB) at their levels
        k, acc = 0, 0
        while k < a.inputs and acc < 32 * MiB:
            acc += int(sizes[k])
            k += 1
        srcs = [pool[int(offs[i]):int(offs[i] + sizes[i])].tobytes() for i in range(k)]
        stock = [bz2.compress(srcs[i], int(levels[i])) for i in range(k)]
        t = time.perf_counter()
        ok = all(bz2.decompress(stock[i]) == srcs[i] for i in range(k))
        dt = time.perf_counter() - t
        res["cpu_baseline"] = {"value": round(acc / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "libbz2",
                               "sample": f"Python bz2 (libbz2 1.0.8) decoding its own streams of the first {k} "
                                         f"inputs ({acc} bytes) at their levels, {dt:.2f} s wall",
                               "identical": ok}
    print(json.dumps(res), flush=True)


def _hip_copy_d2h(ptr, nbytes):
    import numpy as np
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.empty(nbytes, dtype=np.uint8)
    rc = hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 2)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    return out.tobytes()


def _cpu_baseline(data, level, block_bytes, sample_bytes):
    """The oracle (a restatement of the reference path) and stock libbz2
    (Python's bz2) on the host cores, block-parallel over min(nproc, 16)
    threads (the reference's worker cap), on the first sample_bytes of the
    workload; 1 warm-up + median of 5 runs.  Returns the baseline record and
    the oracle's stream of the whole workload (the checker)."""
    import bz2
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    threads = max(1, min(os.cpu_count() or 1, 16))
    n = min(data.size, max(block_bytes, sample_bytes // block_bytes * block_bytes))
    sample = data[:n]
    mo, ro = _median_runs(lambda: oracle.compress(sample, level, block_bytes=block_bytes, threads=threads))
    chunks = [sample[i:i + block_bytes].tobytes() for i in range(0, n, block_bytes)]

    def run_bz2():
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda c: bz2.compress(c, level), chunks))

    mb, rb = _median_runs(run_bz2)
    rec = {"value": round(n / mo / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
           "sample": f"oracle: the first {n} bytes of the workload, block-parallel on {threads} host threads; "
                     f"1 warm-up + median of 5 runs", "runs_s": [round(x, 3) for x in ro]}
    rec.update(_cpu_info())
    rec["libbz2"] = {"value": round(n / mb / 1e9, 4), "unit": "GB/s", "cores": threads,
                     "sample": f"Python bz2 (libbz2 1.0.8) on the same {n} bytes as independent "
                               f"{block_bytes}-byte chunks on {threads} threads; 1 warm-up + median of 5 runs",
                     "runs_s": [round(x, 3) for x in rb]}
    out = oracle.compress(data, level, block_bytes=block_bytes, threads=threads)
    return rec, out


def main():
    a = _args()
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {a.gpus}); runs over several GPUs are not supported",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "batch4096":
        return run_batch(a)
    if a.workload == "serve":
        return run_serve(a)
    if a.workload == "decompress":
        return run_decompress(a)
    if a.workload == "decompress-batch":
        return run_decompress_batch(a)
    import numpy as np
    import torch
    import bzip2gpu_amd as bz

    data, cfg = _workload(a.workload, a.mib)
    level, bb = cfg["level"], cfg["block_bytes"]
    n = data.size
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(data).to(dev)
    torch.cuda.synchronize()

    def batched():
        return bz.compress_batch_device(d_in, [0], [n], [level], bb)

    tune = _autotune(batched, forced=a.tune_force)
    for _ in range(max(0, a.warmup)):
        batched()
    torch.cuda.synchronize()
    if a.calib_copy:
        cp = d_in.clone()
        torch.cuda.synchronize()
        del cp
    bz.probe_kernel(a.probe)
    bz.probe_read()
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        ptr, bo, by = batched()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    probe_ms, probe_n = bz.probe_read()
    bz.probe_kernel(None)
    stages, rounds, rows = bz.last_stage_times()
    ms = wall / a.steps * 1e3
    value = n / (wall / a.steps) / 1e9
    out_batched = _hip_copy_d2h(ptr + int(bo[0]), int(by[0]))

    res = {
        "metric": "input GB/s at 1 MI355X, batched vs plain per-block (bit-exact); p50/p99 latency",
        "value": round(value, 4),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": dict(cfg, path="batched", blocks=(n + bb - 1) // bb),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "stage_ms": {k: round(v, 3) for k, v in stages.items()},
        "bwt_rounds": rounds,
        "compressed_bytes": int(by[0]),
        "autotune": tune,
    }
    rs = bz.last_round_stats()
    # suffix-sort rounds: rows sorted, GPU ms, and GB/s at 20 algorithmic
    # bytes per row per round (read SA 4 + gather 2 ranks 8 + write SA 4 + rank 4)
    res["bwt_round_stats"] = [{"rows": r, "ms": round(t, 3), "gbps_alg": round(20 * r / (t * 1e-3) / 1e9, 1) if t else None}
                              for r, t in rs]
    small_rows, large_rows, pass_rows = bz.last_sort_rows()
    res["sort_rows"] = {"small_groups": small_rows, "large_groups": large_rows, "large_x_passes": pass_rows}
    per_unit, unit = ALG_BYTES.get(a.probe, (0, ""))
    units = {"radix_pass_rows": pass_rows}.get(unit, 0)
    if probe_n and units and per_unit:
        alg = per_unit * units  # per step
        per_launch = alg * a.steps / probe_n
        avg_s = probe_ms / probe_n / 1e3
        achieved = per_launch / avg_s / 1e9
        traffic = None
        if a.traffic and os.path.exists(a.traffic):
            with open(a.traffic) as f:
                traffic = json.load(f).get(f"{cfg['workload']}:{a.probe}")
        res["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                           "frac": round(achieved / 8000.0, 4), "traffic": traffic,
                           "kernel": a.probe, "launches": probe_n, "avg_launch_us": round(probe_ms / probe_n * 1e3, 2),
                           "alg_bytes_per_launch": int(per_launch),
                           "alg_bytes_rule": f"{per_unit} B x {unit}"}
    if not a.no_plain:
        torch.cuda.synchronize()
        t = time.perf_counter()
        out_plain = bz.compress_plain_device(d_in, level, bb)
        dtp = time.perf_counter() - t
        res["plain"] = {"value": round(n / dtp / 1e9, 4), "unit": "GB/s", "seconds": round(dtp, 3),
                        "identical_to_batched": out_plain == out_batched}
        res["batched_over_plain"] = round(value / (n / dtp / 1e9), 2)
    totals = bz.last_batch_totals()
    alg = _alg_bytes(totals)
    pipe = alg / (wall / a.steps) / 1e9
    copy_peak = _copy_peak_gbps()
    res.setdefault("roofline", {})["pipeline"] = {
        "alg_bytes_per_step": int(alg), "achieved": round(pipe, 2), "frac": round(pipe / 8000.0, 4),
        "copy_peak_gbps": round(copy_peak, 1), "frac_of_copy_peak": round(pipe / copy_peak, 4),
        "rule": "per block 2n + 9m + 20 m R + 12 s + W/8, + W/4 assembly", "totals": totals}
    if not a.no_cpu:
        cpu, cpu_out = _cpu_baseline(data, level, bb, a.cpu_sample_mib * MiB)
        res["cpu_baseline"] = cpu
        res["identical_to_oracle"] = cpu_out == out_batched
    elif not a.no_check:
        import oracle
        threads = max(1, min(os.cpu_count() or 1, 16))
        res["identical_to_oracle"] = oracle.compress(data, level, block_bytes=bb, threads=threads) == out_batched
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
