"""Small drivers for rocprofv3 runs (kernel traces and PMC counters).

  python -m tsm_benchmark_amd.profiling atb  --n 33554432 --k 8 --m 8 --dtype f64 --reps 5
  python -m tsm_benchmark_amd.profiling copy --bytes 1073741824 --reps 5
  python -m tsm_benchmark_amd.profiling bwd  --n 8388608 --k 16 --m 16 --dtype f32 --reps 5 [--dA-only]
  python -m tsm_benchmark_amd.profiling batched --dtype f64 --reps 5
  python -m tsm_benchmark_amd.profiling ata  --n 33554432 --k 8 --dtype f64 --reps 5
  python -m tsm_benchmark_amd.profiling tsmm --n 33554432 --k 8 --m 8 --dtype f64 --reps 5

`atb --sets S` uses S input sets in rotation (as bench.py does) and
`--back-to-back` (atb, bwd, batched) launches all reps without a
synchronisation between them.
`batched` is the batched workload of the measurement suite (512 pairs, n
log-uniform in [2^10, 2^21], k, m in {2, 4, 8, 16, 32}) as one plan launch per
rep.

`--dist pow2` draws the entries as sign * 2^e, e uniform in [-20, 20], instead
of uniform on [-1, 1).

`copy` runs torch's device copy of a known byte count: the calibration for the
FETCH_SIZE / WRITE_SIZE counters (bytes counted per byte moved).  Every step
prints one line so a long profiled run shows progress.
"""
from __future__ import annotations

import argparse
import sys

import torch

from . import _lib, ops

DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}
SEED = 20261203


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["atb", "copy", "bwd", "batched", "ata", "tsmm"])
    ap.add_argument("--sets", type=int, default=1)
    ap.add_argument("--back-to-back", action="store_true")
    ap.add_argument("--n", type=int, default=1 << 25)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--m", type=int, default=8)
    ap.add_argument("--dtype", default="f64", choices=list(DT))
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--dist", default="uniform", choices=["uniform", "pow2"])
    ap.add_argument("--dA-only", dest="da_only", action="store_true",
                    help="bwd: the gradient of A only (dB not computed)")
    args = ap.parse_args(argv)
    torch.cuda.set_device(0)
    if args.what == "copy":
        X = torch.ones(args.bytes // 8, dtype=torch.float64, device="cuda")
        Y = torch.empty_like(X)
        for i in range(args.reps):
            Y.copy_(X)
            torch.cuda.synchronize()
            print(f"copy {i}: {args.bytes} bytes read, {args.bytes} bytes written", flush=True)
        return 0
    dt = DT[args.dtype]
    dist = 0 if args.dist == "uniform" else 1
    if args.what == "batched":
        import numpy as np
        rng = np.random.default_rng(SEED)
        shapes = [(int(round(2 ** rng.uniform(10, 21))),) + tuple(int(x) for x in rng.choice([2, 4, 8, 16, 32], 2))
                  for _ in range(512)]
        As = [ops.random_matrix(n, k, dt, seed=SEED, stream_id=3 * i + 1) for i, (n, k, m) in enumerate(shapes)]
        Bs = [ops.random_matrix(n, m, dt, seed=SEED, stream_id=3 * i + 2) for i, (n, k, m) in enumerate(shapes)]
        outs = [torch.empty((k, m), dtype=torch.float64, device="cuda") for n, k, m in shapes]
        byts = sum(As[0].element_size() * n * (k + m) for n, k, m in shapes)
        plan = ops.BatchedPlan(As, Bs, outs)
        print(f"batched {args.dtype}: 512 pairs, {byts} bytes, {plan.items} items, grid {plan.grid}",
              flush=True)
        for i in range(args.reps):
            plan.run()
            if not args.back_to_back:
                torch.cuda.synchronize()
                print(f"batched {i}", flush=True)
        torch.cuda.synchronize()
        print(f"batched: {args.reps} launches done", flush=True)
        plan.close()
        return 0
    sets = [(ops.random_matrix(args.n, args.k, dt, seed=SEED + j, stream_id=1, dist=dist),
             ops.random_matrix(args.n, args.m, dt, seed=SEED + j, stream_id=2, dist=dist))
            for j in range(max(1, args.sets))]
    A, B = sets[0]
    torch.cuda.synchronize()
    if args.what == "atb":
        idx, grid = _lib.plan({"f64": 0, "f32": 1, "bf16": 2}[args.dtype], args.n, args.k, args.m)
        print(f"kernel {_lib.kernels()[idx]['name']} grid {grid}", flush=True)
        C = torch.empty((args.k, args.m), dtype=torch.float64, device="cuda")
        for i in range(args.reps):
            Ai, Bi = sets[i % len(sets)]
            ops.atb(Ai, Bi, out=C)
            if not args.back_to_back:
                torch.cuda.synchronize()
                print(f"atb {i}: {A.element_size() * args.n * (args.k + args.m)} bytes", flush=True)
        torch.cuda.synchronize()
        print(f"atb: {args.reps} launches done", flush=True)
    elif args.what in ("ata", "tsmm"):
        # A^T A reads A only; Y = A X reads A and writes Y (X is k x m)
        if args.what == "ata":
            C = torch.empty((args.k, args.k), dtype=torch.float64, device="cuda")
            fn = lambda: ops.ata(A, out=C)  # noqa: E731
            moved = A.element_size() * args.n * args.k
        else:
            X = torch.randn((args.k, args.m), dtype=torch.float64, device="cuda",
                            generator=torch.Generator(device="cuda").manual_seed(5))
            if dt != torch.float64:
                X = X.float()
            fn = lambda: ops.tsmm(A, X)  # noqa: E731
            moved = A.element_size() * args.n * (args.k + args.m)
        for i in range(args.reps):
            fn()
            if not args.back_to_back:
                torch.cuda.synchronize()
                print(f"{args.what} {i}: {moved} bytes", flush=True)
        torch.cuda.synchronize()
        print(f"{args.what}: {args.reps} launches done", flush=True)
    else:
        G = torch.randn((args.k, args.m), dtype=torch.float32, device="cuda")
        G = G.to(torch.float64 if dt == torch.float64 else torch.float32)
        for i in range(args.reps):
            ops.atb_backward(A, B, G, need_dB=not args.da_only)
            if not args.back_to_back:
                torch.cuda.synchronize()
                print(f"bwd {i}", flush=True)
        torch.cuda.synchronize()
        print(f"bwd: {args.reps} launches done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
