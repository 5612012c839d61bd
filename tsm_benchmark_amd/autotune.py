"""Autotuner for the generated C = A^T B kernels (runs on the GPU).

For each tuned shape (dtype, k class, m class) it times every compiled
variant of that class (register tile, rows per tile, pipeline depth, load
policy) at several persistent-grid sizes on a problem of `--gib` GiB of A + B,
checks that the variant's result agrees with the default kernel's, and
writes the fastest configuration per class to the tile table
(tuning/gfx950.json).  The generator bakes that table into the library as
the tuned defaults:

    python -m tsm_benchmark_amd.autotune --out tsm_benchmark_amd/tuning/gfx950.json
    python -m tsm_benchmark_amd.build
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import torch

from . import _lib, codegen, ops

SEED = 20261203
DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}
DTYPE_ID = {"f64": 0, "f32": 1, "bf16": 2}
GRID_MULTS = (1, 2, 3, 4, 6, 8)


def time_call(fn, reps: int) -> float:
    """Seconds per call of `reps` back-to-back calls after 3 warmup calls
    (HIP events on the current stream): the steady-state rate of a caller
    that issues the product repeatedly."""
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / reps


def settle(fn, seconds: float = 0.3) -> None:
    """Back-to-back calls for `seconds` before the first timing of a shape.
    The shader clock drops about 1 ms after a streaming load starts on an
    idle chip and takes some 20 ms to recover; without this the variants
    timed first would be measured in that dip."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(8):
            fn()
        torch.cuda.synchronize()


def accurate_on_pow2(dtype: str, kc: int, mc: int, cfg) -> bool:
    """FP32/BF16 storage: the Frobenius error against the FP64 kernel on the
    same rounded sign * 2^e inputs (entries over 41 binades) is <= 1e-6."""
    if dtype == "f64":
        return True
    n = 1 << 20
    A = ops.random_matrix(n, kc, DT[dtype], seed=SEED, stream_id=11, dist=_lib.DIST_POW2)
    B = ops.random_matrix(n, mc, DT[dtype], seed=SEED, stream_id=12, dist=_lib.DIST_POW2)
    got = ops.atb(A, B, out_dtype=torch.float64, config=cfg)
    ref = ops.atb(A.double(), B.double())
    return ((got - ref).norm() / ref.norm()).item() <= 1e-6


def tune_shape(dtype: str, kc: int, mc: int, gib: float, reps: int, log,
               mults: tuple = GRID_MULTS) -> dict:
    es = codegen.ELEM_BYTES[dtype]
    n = int(gib * (1 << 30) / (es * (kc + mc)))
    A = ops.random_matrix(n, kc, DT[dtype], seed=SEED, stream_id=1)
    B = ops.random_matrix(n, mc, DT[dtype], seed=SEED, stream_id=2)
    C = torch.empty((kc, mc), dtype=torch.float64, device="cuda")
    ref = ops.atb(A, B).double()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    nbytes = es * n * (kc + mc)
    results = []
    settle(lambda: ops.atb(A, B, out=C))
    default_t = time_call(lambda: ops.atb(A, B, out=C), reps)
    for kd in _lib.kernels():
        if kd["in_type"] != DTYPE_ID[dtype] or kd["kc"] != kc or kd["mc"] != mc:
            continue
        # the direct engine streams each wave's own rows: no tile order to choose
        orders = (0,) if kd["engine"] == _lib.ENGINE_DIRECT else (0, 1)
        for mult, il in [(mu, il) for mu in mults for il in orders]:
            cfg = _lib.config_of(kd, wgs_per_cu=mult, interleave=il)
            t = time_call(lambda: ops.atb(A, B, out=C, config=cfg), reps)
            rel = ((C - ref).norm() / ref.norm()).item()
            tol = 1e-13 if dtype == "f64" else 1e-6
            if not rel <= tol:
                raise RuntimeError(f"{kd['name']} x{mult}: result differs from the default "
                                   f"kernel by {rel:.3g}")
            log(f"{kd['name']:40s} wgs/cu {mult} il {il}: {t * 1e6:9.1f} us "
                f"{nbytes / t / 1e9:8.1f} GB/s")
            results.append((t, kd["name"], kd, mult, il))
    results.sort(key=lambda r: (r[0], r[1]))

    def fastest(direct_ok: bool):
        for t, _, kd, mult, il in results:
            if not direct_ok and kd["engine"] == _lib.ENGINE_DIRECT:
                continue
            if accurate_on_pow2(dtype, kc, mc, _lib.config_of(kd, wgs_per_cu=mult, interleave=il)):
                return t, kd, mult, il
            log(f"{kd['name']} misses 1e-6 on sign * 2^e inputs: skipped")
        raise RuntimeError(f"no {dtype} {kc}x{mc} kernel meets 1e-6 on sign * 2^e inputs")

    t, kd, mult, il = fastest(True)
    entry = {"dtype": dtype, "kc": kc, "mc": mc, "tm": kd["tm"], "tn": kd["tn"],
             "rows": kd["rows"], "pf": kd["pf"], "nt": kd["nt"], "wgs_per_cu": mult,
             "interleave": il,
             "gbps": round(nbytes / t / 1e9, 1), "default_gbps": round(nbytes / default_t / 1e9, 1),
             "n": n, "cus": cus}
    if kd["engine"] == _lib.ENGINE_DIRECT:
        # the staged kernel for inputs the direct engine cannot read
        entry["direct"] = 1
        ft, fkd, _, fil = fastest(False)
        entry["fallback"] = {"tm": fkd["tm"], "tn": fkd["tn"], "rows": fkd["rows"],
                             "pf": fkd["pf"], "nt": fkd["nt"], "interleave": fil,
                             "gbps": round(nbytes / ft / 1e9, 1)}
    del A, B
    torch.cuda.empty_cache()
    return entry


def tune_sym(dtype: str, kc: int, gib: float, reps: int, log) -> dict:
    """A^T A of one square class: every compiled variant x grid x tile order,
    on A of `gib` GiB."""
    es = codegen.ELEM_BYTES[dtype]
    n = int(gib * (1 << 30) / (es * kc))
    A = ops.random_matrix(n, kc, DT[dtype], seed=SEED, stream_id=1)
    C = torch.empty((kc, kc), dtype=torch.float64, device="cuda")
    ref = ops.ata(A).double()
    _, defaults = codegen.library_kernels(codegen.load_tuning())
    variants = [c for c in codegen.sym_kernels(defaults, codegen.load_tuning())[0]
                if c.dtype == dtype and c.kc == kc]
    nbytes = es * n * kc
    results = []
    settle(lambda: ops.ata(A, out=C))
    default_t = time_call(lambda: ops.ata(A, out=C), reps)
    for c in variants:
        orders = (0,) if c.direct else (0, 1)
        for mult, il in [(mu, il) for mu in GRID_MULTS for il in orders]:
            cfg = _lib.make_config(tm=c.tm, tn=c.tn, rows=c.rows, pf=c.pf, nt=c.nt,
                                   wgs_per_cu=mult, interleave=il, engine=c.engine)
            t = time_call(lambda: ops.ata(A, out=C, config=cfg), reps)
            rel = ((C - ref).norm() / ref.norm()).item()
            tol = 1e-13 if dtype == "f64" else 1e-6
            if not rel <= tol:
                raise RuntimeError(f"{c.name} x{mult}: result differs from the default by {rel:.3g}")
            log(f"{c.name:40s} wgs/cu {mult} il {il}: {t * 1e6:9.1f} us "
                f"{nbytes / t / 1e9:8.1f} GB/s")
            results.append((t, c.name, c, mult, il))
    results.sort(key=lambda r: (r[0], r[1]))
    t, _, c, mult, il = results[0]
    del A
    torch.cuda.empty_cache()
    return {"dtype": dtype, "kc": kc, "tm": c.tm, "tn": c.tn, "rows": c.rows, "pf": c.pf,
            "nt": c.nt, "wgs_per_cu": mult, "interleave": il, "direct": c.direct,
            "gbps": round(nbytes / t / 1e9, 1), "default_gbps": round(nbytes / default_t / 1e9, 1),
            "n": n}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=codegen.TUNING_JSON)
    ap.add_argument("--gib", type=float, default=2.0, help="GiB of A + B per problem")
    ap.add_argument("--reps", type=int, default=7)
    ap.add_argument("--shapes", nargs="*", default=None,
                    help="dtype:kcxmc entries (default: the generator's tuned shapes)")
    ap.add_argument("--ops", nargs="*", default=["atb", "ata"], choices=["atb", "ata"],
                    help="tune the A^T B kernels, the A^T A kernels, or both")
    ap.add_argument("--all-classes", action="store_true",
                    help="also every other width class (its compiled kernels, 1 or 2 "
                         "workgroups per CU, both tile orders)")
    args = ap.parse_args(argv)
    torch.cuda.set_device(0)
    shapes = []
    if args.shapes:
        for s in args.shapes:
            dt, km = s.split(":")
            k, m = km.split("x")
            shapes.append((dt, int(k), int(m)))
    else:
        for dt, lst in codegen.TUNE_SHAPES.items():
            shapes += [(dt, k, m) for k, m in lst]
    table = codegen.load_tuning(args.out) if os.path.exists(args.out) else {"entries": []}
    entries = {(e["dtype"], e["kc"], e["mc"]): e for e in table.get("entries", [])}
    sym_entries = {(e["dtype"], e["kc"]): e for e in table.get("sym_entries", [])}
    t0 = time.time()
    log = lambda s: print(s, flush=True)  # noqa: E731
    if "atb" in args.ops:
        for dt, k, m in shapes:
            print(f"== {dt} {k}x{m}", flush=True)
            e = tune_shape(dt, k, m, args.gib, args.reps, log)
            print(f"-> {e}", flush=True)
            entries[(dt, k, m)] = e
        if args.all_classes:
            for dt in codegen.DTYPES:
                for k in codegen.CLASSES:
                    for m in codegen.CLASSES:
                        if (dt, k, m) in shapes or (k, m) in codegen.TUNE_SHAPES[dt]:
                            continue
                        print(f"== {dt} {k}x{m} (grid and order)", flush=True)
                        e = tune_shape(dt, k, m, args.gib, args.reps, log, mults=(1, 2))
                        print(f"-> {e}", flush=True)
                        entries[(dt, k, m)] = e
    if "ata" in args.ops:
        for dt, k, m in shapes:
            if k != m:
                continue
            print(f"== A^T A {dt} {k}x{k}", flush=True)
            e = tune_sym(dt, k, args.gib, args.reps, log)
            print(f"-> {e}", flush=True)
            sym_entries[(dt, k)] = e
    out = {"format": codegen.TUNING_FORMAT,
           "device": torch.cuda.get_device_properties(0).gcnArchName,
           "rocm": getattr(torch.version, "hip", None),
           "library_version": _lib.lib().tsm_version(),
           "generated_by": "python -m tsm_benchmark_amd.autotune",
           "entries": sorted(entries.values(), key=lambda e: (e["dtype"], e["kc"], e["mc"])),
           "sym_entries": sorted(sym_entries.values(), key=lambda e: (e["dtype"], e["kc"]))}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {args.out} ({len(out['entries'])} entries) in {time.time() - t0:.0f}s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
