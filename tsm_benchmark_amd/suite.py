"""Measurement suite over the five workloads of the tall-skinny AᵀB path.

  python -m tsm_benchmark_amd.suite [--quick] [--out report.json]

  plain       FP64 k = m = 8, n = 2^22                        GB/s, % of 8 TB/s
  sweep       FP64 k = m in {1..64}, (4,32), (32,4), A+B = 4 GiB, uniform and
              sign*2^e inputs                                 GB/s, % HBM, % FP64 peak
  precision   k = m = 32, n = 2^25: FP32 and BF16 storage vs FP64; error of
              each against the FP64 kernel on the SAME rounded inputs
  backward    FP32 n = 2^23, k = m = 16: forward, backward, (fwd+bwd)/fwd
  batched     512 pairs, n log-uniform in [2^10, 2^21], k, m in {2..32}, FP64
              and FP32: one persistent launch vs one launch per pair
  siblings    C = A^T A (FP64, A = 2 GiB, k = 8 and 32) against tsm_atb(A, A);
              Y = A X (FP64, k = m = 8, n = 2^25)
  matrix      FP64 and FP32 k = m = 64, A+B = 4 GiB, both distributions: the
              matrix-core bound classes, GB/s and GFLOP/s with the fraction
              of the dtype's dense matrix peak

Times are per call over back-to-back calls after at least 0.1 s of warmup
calls (HIP events on the launch stream), inputs resident in HBM: the
steady-state rate, as in bench.py and the autotuner.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import _lib, ops

SEED = 20261203
HBM = 8.0e12
FP64_PEAK = 78.6e12
FP32_PEAK = 157.3e12
DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}


def timed(fn, reps: int = 10, warm_s: float = 0.1) -> float:
    """Seconds per call over `reps` back-to-back calls, after at least three
    warmup calls and `warm_s` seconds of them: right after its inputs are
    allocated a pass over a few GiB runs up to 25% slower for the first
    several milliseconds."""
    t0 = time.perf_counter()
    calls = 0
    while calls < 3 or time.perf_counter() - t0 < warm_s:
        fn()
        torch.cuda.synchronize()
        calls += 1
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / reps


def rand(n, w, dt, sid, dist=0):
    return ops.random_matrix(n, w, DT[dt], seed=SEED, stream_id=sid, dist=dist)


def plain(reps):
    n, k, m = 1 << 22, 8, 8
    A, B = rand(n, k, "f64", 1), rand(n, m, "f64", 2)
    C = torch.empty((k, m), dtype=torch.float64, device="cuda")
    t = timed(lambda: ops.atb(A, B, out=C), reps)
    gb = (8 * n * (k + m) + 8 * k * m) / t / 1e9
    return {"n": n, "k": k, "m": m, "us": round(t * 1e6, 1), "GBps": round(gb, 1),
            "hbm_frac": round(gb * 1e9 / HBM, 4)}


def sweep(reps, quick):
    shapes = [(1, 1), (2, 2), (4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 32), (32, 4)]
    if quick:
        shapes = [(1, 1), (8, 8), (64, 64), (4, 32)]
    out = []
    for k, m in shapes:
        n = (1 << 32) // (8 * (k + m)) if (k, m) not in ((4, 32), (32, 4)) else 14_913_080
        for dist in (0, 1):
            A, B = rand(n, k, "f64", 1, dist), rand(n, m, "f64", 2, dist)
            C = torch.empty((k, m), dtype=torch.float64, device="cuda")
            t = timed(lambda: ops.atb(A, B, out=C), reps)
            byts = 8 * n * (k + m) + 8 * k * m
            out.append({"k": k, "m": m, "n": n, "dist": ["uniform", "pow2"][dist],
                        "us": round(t * 1e6, 1), "GBps": round(byts / t / 1e9, 1),
                        "hbm_frac": round(byts / t / HBM, 4),
                        "fp64_frac": round(2 * n * k * m / t / FP64_PEAK, 4)})
            del A, B
    return out


def precision(reps, quick):
    n, k, m = (1 << 23 if quick else 1 << 25), 32, 32
    res = {"n": n, "k": k, "m": m}
    for dist in (0, 1):
        A64, B64 = rand(n, k, "f64", 1, dist), rand(n, m, "f64", 2, dist)
        C = torch.empty((k, m), dtype=torch.float64, device="cuda")
        t64 = timed(lambda: ops.atb(A64, B64, out=C), reps)
        row = {"fp64_us": round(t64 * 1e6, 1)}
        for dt in ("f32", "bf16"):
            A, B = rand(n, k, dt, 1, dist), rand(n, m, dt, 2, dist)
            t = timed(lambda: ops.atb(A, B, out=C), reps)
            got = ops.atb(A, B, out_dtype=torch.float64)
            ref = ops.atb(A.double(), B.double())  # FP64 kernel on the same rounded inputs
            absC = ops.atb(A.double().abs(), B.double().abs())
            row[dt] = {"us": round(t * 1e6, 1),
                       "GBps": round(A.element_size() * n * (k + m) / t / 1e9, 1),
                       "speedup_vs_fp64": round(t64 / t, 3),
                       "frob_rel_err": float(((got - ref).norm() / ref.norm()).item()),
                       "max_err_over_absAB": float(((got - ref).abs() / absC).max().item())}
            del A, B
        res[["uniform", "pow2"][dist]] = row
        del A64, B64
    return res


def backward(reps):
    n, k, m = 1 << 23, 16, 16
    A, B = rand(n, k, "f32", 1), rand(n, m, "f32", 2)
    G = torch.randn((k, m), dtype=torch.float32, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(5))
    C = torch.empty((k, m), dtype=torch.float32, device="cuda")
    tf = timed(lambda: ops.atb(A, B, out=C), reps)
    tb = timed(lambda: ops.atb_backward(A, B, G), reps)
    tb_a = timed(lambda: ops.atb_backward(A, B, G, need_dB=False), reps)
    fwd_bytes = 4 * n * (k + m)
    return {"n": n, "k": k, "m": m, "fwd_us": round(tf * 1e6, 1), "bwd_us": round(tb * 1e6, 1),
            "bwd_dA_only_us": round(tb_a * 1e6, 1),
            "fwd_GBps": round(fwd_bytes / tf / 1e9, 1),
            "bwd_GBps": round(2 * fwd_bytes / tb / 1e9, 1),
            "fwd_plus_bwd_over_fwd": round((tf + tb) / tf, 3),
            "byte_bound_ratio": 3.0}


def batched(reps):
    rng = np.random.default_rng(SEED)
    shapes = [(int(round(2 ** rng.uniform(10, 21))),) + tuple(int(x) for x in rng.choice([2, 4, 8, 16, 32], 2))
              for _ in range(512)]
    res = {"pairs": 512}
    for dt in ("f64", "f32"):
        As = [rand(n, k, dt, 3 * i + 1) for i, (n, k, m) in enumerate(shapes)]
        Bs = [rand(n, m, dt, 3 * i + 2) for i, (n, k, m) in enumerate(shapes)]
        outs = [torch.empty((k, m), dtype=torch.float64, device="cuda") for n, k, m in shapes]
        es = As[0].element_size()
        byts = sum(es * n * (k + m) for n, k, m in shapes)
        plan = ops.BatchedPlan(As, Bs, outs)
        t_plan = timed(plan.run, reps)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            plan.run()
        t_graph = timed(g.replay, reps)
        t_call = timed(lambda: ops.atb_batched(As, Bs, outs=outs), reps)

        def per_pair():
            for A, B, C in zip(As, Bs, outs):
                ops.atb(A, B, out=C)
        t_pp = timed(per_pair, max(3, reps // 3))
        res[dt] = {"bytes": byts, "items": plan.items, "grid": plan.grid,
                   "plan_us": round(t_plan * 1e6, 1), "plan_GBps": round(byts / t_plan / 1e9, 1),
                   "graph_us": round(t_graph * 1e6, 1), "graph_GBps": round(byts / t_graph / 1e9, 1),
                   "call_us": round(t_call * 1e6, 1),
                   "per_pair_us": round(t_pp * 1e6, 1), "per_pair_GBps": round(byts / t_pp / 1e9, 1)}
        plan.close()
        del As, Bs
    return res


def siblings(reps):
    out = {}
    for k in (8, 32):
        n = (1 << 31) // (8 * k)
        A = rand(n, k, "f64", 1)
        C = torch.empty((k, k), dtype=torch.float64, device="cuda")
        t_ata = timed(lambda: ops.ata(A, out=C), reps)
        t_atb = timed(lambda: ops.atb(A, A, out=C), reps)
        byts = 8 * n * k + 8 * k * k
        out[f"ata_k{k}"] = {"n": n, "k": k, "ata_us": round(t_ata * 1e6, 1),
                            "ata_GBps": round(byts / t_ata / 1e9, 1),
                            "atb_AA_us": round(t_atb * 1e6, 1),
                            "speedup_vs_atb_AA": round(t_atb / t_ata, 3)}
        del A
    n, k, m = 1 << 25, 8, 8
    A = rand(n, k, "f64", 1)
    X = torch.randn((k, m), dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(5))
    t = timed(lambda: ops.tsmm(A, X), reps)
    byts = 8 * n * (k + m) + 8 * k * m
    out["tsmm_k8_m8"] = {"n": n, "k": k, "m": m, "us": round(t * 1e6, 1),
                         "GBps": round(byts / t / 1e9, 1)}
    return out


def matrix(reps):
    out = {}
    for dt, es, peak in (("f64", 8, FP64_PEAK), ("f32", 4, FP32_PEAK)):
        k = m = 64
        n = (1 << 32) // (es * (k + m))
        idx, _ = _lib.plan({"f64": _lib.F64, "f32": _lib.F32}[dt], n, k, m)
        row = {"n": n, "k": k, "m": m, "kernel": _lib.kernels()[idx]["name"]}
        for dist in (0, 1):
            A, B = rand(n, k, dt, 1, dist), rand(n, m, dt, 2, dist)
            C = torch.empty((k, m), dtype=torch.float64, device="cuda")
            t = timed(lambda: ops.atb(A, B, out=C), reps)
            flops = 2 * n * k * m
            row[["uniform", "pow2"][dist]] = {
                "us": round(t * 1e6, 1), "GBps": round(es * n * (k + m) / t / 1e9, 1),
                "GFLOPs": round(flops / t / 1e9, 1), "matrix_peak_frac": round(flops / t / peak, 4)}
            del A, B
        out[dt] = row
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--only", nargs="*", default=None)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    torch.cuda.set_device(0)
    parts = {"plain": lambda: plain(args.reps), "sweep": lambda: sweep(args.reps, args.quick),
             "precision": lambda: precision(args.reps, args.quick),
             "backward": lambda: backward(args.reps), "batched": lambda: batched(args.reps),
             "siblings": lambda: siblings(args.reps), "matrix": lambda: matrix(args.reps)}
    report = {}
    for name, fn in parts.items():
        if args.only and name not in args.only:
            continue
        report[name] = fn()
        torch.cuda.empty_cache()
        print(json.dumps({name: report[name]}), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
