"""CPU oracle for the tall-skinny product C = A^T B (numpy front end).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this package, and only as the
checker (or, for the baseline, as the timed CPU restatement of the reference
kernels).  The product path never imports it.

The C code lives in ``oracle.c`` (inputs, rounding, Dot2, exact int128, the
reference's naive loop and acceptance rule, backward) and ``ref_kernels.cc``
(the reference's SYCL kernels restated for host cores).  ``make -C oracle``
builds ``liboracle.so``.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

SEED = 20261203  # the global seed of the synthetic workloads

DIST_UNIFORM, DIST_POW2, DIST_NORMAL = 0, 1, 2

# Variant ids of the restated reference kernels (oracle.h).
REF_GEN_BASIC, REF_GEN_LF, REF_GEN_SG, REF_GEN_SG_LF = 0, 1, 2, 3
REF_V1, REF_V2, REF_V3, REF_V4, REF_V5, REF_V6, REF_V7, REF_V8, REF_V9 = (
    11, 12, 13, 14, 15, 16, 17, 18, 19)


class RefCfg(ct.Structure):
    _fields_ = [
        ("variant", ct.c_int), ("tm", ct.c_int), ("tn", ct.c_int),
        ("stride", ct.c_int64), ("local_size", ct.c_int),
        ("sg_size_gen", ct.c_int), ("sg_size_rt", ct.c_int),
        ("unroll", ct.c_int), ("worker_mult", ct.c_int), ("num_cu", ct.c_int),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        P, I64, I, D, U64 = ct.c_void_p, ct.c_int64, ct.c_int, ct.c_double, ct.c_uint64
        L.oracle_rng_word.restype = U64
        L.oracle_rng_word.argtypes = [U64, U64, U64]
        L.oracle_rng_value.restype = D
        L.oracle_rng_value.argtypes = [I, U64, U64, U64]
        L.oracle_fill_f64.argtypes = [I, U64, U64, I64, I64, P, I64]
        L.oracle_round_bf16.restype = ct.c_uint16
        L.oracle_round_bf16.argtypes = [D]
        L.oracle_round_matrix_f32.argtypes = [I64, P, P]
        L.oracle_round_matrix_bf16.argtypes = [I64, P, P]
        L.oracle_atb_dot2.argtypes = [I64, I, I, P, I64, P, I64, P, I64, P]
        L.oracle_atb_exact.restype = I
        L.oracle_atb_exact.argtypes = [I64, I, I, P, I64, I, P, I64, I, P, I64]
        L.oracle_atb_naive.argtypes = [I64, I, I, P, I64, P, I64, P, I64]
        L.oracle_first_mismatch.restype = I64
        L.oracle_first_mismatch.argtypes = [I64, P, P, D, D]
        L.oracle_atb_backward.argtypes = [I64, I, I, P, I64, P, I64, P, I64, P, I64, P, I64]
        L.oracle_ref_run.restype = I
        L.oracle_ref_run.argtypes = [ct.POINTER(RefCfg), I64, I, I, P, P, P]
        L.oracle_suggest_stride.restype = I64
        L.oracle_suggest_stride.argtypes = [P, I, P, I, I, I, I]
        L.oracle_validate_stride.restype = I
        L.oracle_validate_stride.argtypes = [P, I, P, I, I64, I, I]
        L.oracle_ref_time.restype = I
        L.oracle_ref_time.argtypes = [ct.POINTER(RefCfg), I64, I, I, P, P, P, I, I, D, P]
        L.oracle_num_threads.restype = I
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _f64c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- inputs ----------------------------------------------------------------
def rng_word(seed: int, stream: int, idx: int) -> int:
    return lib().oracle_rng_word(seed, stream, idx)


def gen_matrix(rows: int, cols: int, *, dist: int = DIST_UNIFORM, seed: int = SEED,
               stream: int = 0) -> np.ndarray:
    """rows x cols FP64 matrix of the counter-based synthetic distribution."""
    out = np.empty((rows, cols), dtype=np.float64)
    lib().oracle_fill_f64(dist, seed, stream, rows, cols, _ptr(out), cols)
    return out


def round_f32(x: np.ndarray) -> np.ndarray:
    x = _f64c(x)
    out = np.empty(x.shape, dtype=np.float32)
    lib().oracle_round_matrix_f32(x.size, _ptr(x), _ptr(out))
    return out


def round_bf16_bits(x: np.ndarray) -> np.ndarray:
    """Single round-to-nearest-even from FP64 to bf16; returns uint16 patterns."""
    x = _f64c(x)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().oracle_round_matrix_bf16(x.size, _ptr(x), _ptr(out))
    return out


def bf16_bits_to_f64(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


# ---- oracles -----------------------------------------------------------------
def atb_dot2(A, B, *, with_abs: bool = False):
    """Dot2-compensated C = A^T B in FP64 (optionally also |A|^T |B|)."""
    A, B = _f64c(A), _f64c(B)
    n, k = A.shape
    m = B.shape[1]
    assert B.shape[0] == n
    C = np.empty((k, m), dtype=np.float64)
    absC = np.empty((k, m), dtype=np.float64) if with_abs else None
    lib().oracle_atb_dot2(n, k, m, _ptr(A), k, _ptr(B), m, _ptr(C), m,
                          _ptr(absC) if with_abs else None)
    return (C, absC) if with_abs else C


def atb_exact(A, B, qa: int, qb: int) -> np.ndarray:
    """Exact (correctly rounded) C = A^T B for entries that are multiples of 2^-q."""
    A, B = _f64c(A), _f64c(B)
    n, k = A.shape
    m = B.shape[1]
    C = np.empty((k, m), dtype=np.float64)
    rc = lib().oracle_atb_exact(n, k, m, _ptr(A), k, qa, _ptr(B), m, qb, _ptr(C), m)
    if rc != 0:
        raise ValueError(f"oracle_atb_exact failed ({rc}): inputs not on the 2^-q grid or overflow")
    return C


def atb_naive(A, B) -> np.ndarray:
    """The reference's sequential host loop (main.cc:14-27)."""
    A, B = _f64c(A), _f64c(B)
    n, k = A.shape
    m = B.shape[1]
    C = np.empty((k, m), dtype=np.float64)
    lib().oracle_atb_naive(n, k, m, _ptr(A), k, _ptr(B), m, _ptr(C), m)
    return C


def first_mismatch(ref, got, rel_tol: float = 1e-6, abs_tol: float = 1e-10) -> int:
    """The reference's acceptance rule (main.cc:29-41); -1 when all entries pass."""
    ref, got = _f64c(ref).ravel(), _f64c(got).ravel()
    return int(lib().oracle_first_mismatch(ref.size, _ptr(ref), _ptr(got), rel_tol, abs_tol))


def atb_backward(A, B, G):
    """dA = B G^T and dB = A G for L = <G, A^T B> (Dot2 per entry)."""
    A, B, G = _f64c(A), _f64c(B), _f64c(G)
    n, k = A.shape
    m = B.shape[1]
    dA = np.empty((n, k), dtype=np.float64)
    dB = np.empty((n, m), dtype=np.float64)
    lib().oracle_atb_backward(n, k, m, _ptr(A), k, _ptr(B), m, _ptr(G), m,
                              _ptr(dA), k, _ptr(dB), m)
    return dA, dB


# ---- restated reference kernels -------------------------------------------------
def ref_cfg(variant: int, *, tm: int = 4, tn: int | None = None, stride: int = 524288,
            local_size: int = 256, sg_size_gen: int = 32, sg_size_rt: int | None = None,
            unroll: int = 1, worker_mult: int = 64, num_cu: int = 256) -> RefCfg:
    return RefCfg(variant, tm, tm if tn is None else tn, stride, local_size, sg_size_gen,
                  sg_size_gen if sg_size_rt is None else sg_size_rt, unroll, worker_mult, num_cu)


def ref_run(cfg: RefCfg, A, B, C=None) -> np.ndarray:
    """C += A^T B with a restated reference kernel (C zeroed when not given)."""
    A, B = _f64c(A), _f64c(B)
    n, k = A.shape
    m = B.shape[1]
    if C is None:
        C = np.zeros((k, m), dtype=np.float64)
    rc = lib().oracle_ref_run(ct.byref(cfg), n, k, m, _ptr(A), _ptr(B), _ptr(C))
    if rc != 0:
        raise ValueError("unsupported reference-kernel configuration")
    return C


def suggest_stride(Ms, tiles, local_size=256, sg_size=32, target_wgs=2048) -> int:
    Ms = np.asarray(Ms, dtype=np.int32)
    tiles = np.asarray(tiles, dtype=np.int32)
    return int(lib().oracle_suggest_stride(_ptr(Ms), Ms.size, _ptr(tiles), tiles.size,
                                           local_size, sg_size, target_wgs))


def validate_stride(Ms, tiles, stride, local_size=256, sg_size=32) -> bool:
    Ms = np.asarray(Ms, dtype=np.int32)
    tiles = np.asarray(tiles, dtype=np.int32)
    return bool(lib().oracle_validate_stride(_ptr(Ms), Ms.size, _ptr(tiles), tiles.size,
                                             stride, local_size, sg_size))


def ref_time(cfg: RefCfg, A, B, *, warmup: int = 10, runs: int = 30,
             budget_s: float = 20.0) -> np.ndarray:
    """Wall times (ns) of `runs` restated-reference-kernel calls (C zeroed outside)."""
    A, B = _f64c(A), _f64c(B)
    n, k = A.shape
    m = B.shape[1]
    C = np.zeros((k, m), dtype=np.float64)
    times = np.zeros(runs, dtype=np.float64)
    done = lib().oracle_ref_time(ct.byref(cfg), n, k, m, _ptr(A), _ptr(B), _ptr(C),
                                 warmup, runs, budget_s, _ptr(times))
    if done < 0:
        raise ValueError("unsupported reference-kernel configuration")
    return times[:done]


def num_threads() -> int:
    return int(lib().oracle_num_threads())
