#!/usr/bin/env python3
"""Writes tests/golden/atb_golden.json: exact C = A^T B for small synthetic
problems, from the exact int128 oracle on the counter-based inputs (A: stream
1, B: stream 2, seed 20261203), with |A|^T |B| for the elementwise bound.
FP32 / BF16 cases hold the exact product of the ROUNDED inputs.

The reference ships no golden vectors; these fix the oracle's outputs so that
a later change to the oracle or to the generator shows up as a diff.
Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

CASES = [  # (storage, dist, n, k, m)
    ("f64", 0, 1, 1, 1), ("f64", 0, 1000, 3, 5), ("f64", 0, 4096, 8, 8),
    ("f64", 1, 4096, 4, 4), ("f64", 0, 65537, 4, 32), ("f64", 0, 65536, 32, 4),
    ("f64", 0, 16384, 64, 64), ("f64", 1, 100000, 16, 16), ("f64", 0, 262144, 2, 2),
    ("f32", 0, 8192, 32, 32), ("bf16", 0, 8192, 32, 32), ("f32", 1, 4096, 16, 16),
    ("bf16", 1, 4096, 8, 8),
]


def inputs(storage, dist, n, k, m):
    A = oracle.gen_matrix(n, k, dist=dist, seed=oracle.SEED, stream=1)
    B = oracle.gen_matrix(n, m, dist=dist, seed=oracle.SEED, stream=2)
    if storage == "f32":
        A, B = oracle.round_f32(A).astype("float64"), oracle.round_f32(B).astype("float64")
    elif storage == "bf16":
        A = oracle.bf16_bits_to_f64(oracle.round_bf16_bits(A))
        B = oracle.bf16_bits_to_f64(oracle.round_bf16_bits(B))
    return A, B


def main():
    out = []
    for storage, dist, n, k, m in CASES:
        A, B = inputs(storage, dist, n, k, m)
        q = 52 if dist == 0 else 20
        C = oracle.atb_exact(A, B, q, q)
        _, absC = oracle.atb_dot2(A, B, with_abs=True)
        out.append({"storage": storage, "dist": dist, "seed": oracle.SEED, "n": n, "k": k,
                    "m": m, "C": [float(x).hex() for x in C.ravel()],
                    "absC": [float(x).hex() for x in absC.ravel()]})
    with open(os.path.join(HERE, "atb_golden.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": out}, f)


if __name__ == "__main__":
    main()
