"""Pins the CPU oracle against checks that do not restate it.

* the input generator: the published SplitMix64 outputs;
* the exact int128 oracle: rational arithmetic (fractions.Fraction) on tiny
  inputs and math.fsum (correctly rounded) of exact TwoProduct splits;
* Dot2: the exact oracle, within one unit in the last place;
* the reference's naive loop (main.cc:14-27): left-to-right Python sums;
* the reference's acceptance rule (main.cc:29-41): planted pass/fail cases;
* storage rounding: numpy's float32 cast and bf16 known answers, including a
  value where FP64->FP32->BF16 double rounding differs from one rounding;
* the backward closed form: numpy matmul and FP64 central differences;
* invariants: transposition, linearity, split-n additivity, 2^j scaling.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

# Published SplitMix64 outputs for state 0 (Vigna's reference generator):
SPLITMIX64_SEED0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_splitmix64_known_answers(orc):
    assert [orc.rng_word(0, 0, i) for i in range(3)] == SPLITMIX64_SEED0


def test_uniform_and_pow2_distributions(orc):
    A = orc.gen_matrix(4096, 3, dist=orc.DIST_UNIFORM, seed=7, stream=1)
    assert A.min() >= -1.0 and A.max() < 1.0
    scaled = A * 2.0**52
    assert np.all(scaled == np.round(scaled)), "uniform values are multiples of 2^-52"
    w = orc.rng_word(7, 1, 5)
    assert Fraction(A.flat[5]) == Fraction(w >> 11, 2**52) - 1
    P = orc.gen_matrix(4096, 3, dist=orc.DIST_POW2, seed=7, stream=2)
    e = np.log2(np.abs(P))
    assert np.all(e == np.round(e)) and e.min() == -20 and e.max() == 20
    assert 0.45 < np.mean(P > 0) < 0.55
    # values depend on (seed, stream, index) only
    assert not np.array_equal(A, orc.gen_matrix(4096, 3, seed=7, stream=2))
    assert np.array_equal(orc.gen_matrix(10, 3, seed=7, stream=1), A[:10])


def _dekker_products(a, b):
    """Exact products a*b = p + e without FMA (Dekker/Veltkamp splitting)."""
    split = 134217729.0
    p = a * b
    ca, cb = split * a, split * b
    ah = ca - (ca - a)
    al = a - ah
    bh = cb - (cb - b)
    bl = b - bh
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
    return p, e


def _fsum_atb(A, B):
    k, m = A.shape[1], B.shape[1]
    C = np.empty((k, m))
    for i in range(k):
        for j in range(m):
            p, e = _dekker_products(A[:, i], B[:, j])
            C[i, j] = math.fsum(np.concatenate([p, e]))
    return C


def test_exact_oracle_vs_rationals(orc):
    A = orc.gen_matrix(37, 3, seed=11, stream=1)
    B = orc.gen_matrix(37, 2, seed=11, stream=2)
    C = orc.atb_exact(A, B, 52, 52)
    for i in range(3):
        for j in range(2):
            exact = sum(Fraction(A[r, i]) * Fraction(B[r, j]) for r in range(37))
            assert C[i, j] == float(exact)  # float(Fraction) rounds to nearest even


@pytest.mark.parametrize("dist,q", [(0, 52), (1, 20)])
def test_exact_oracle_vs_fsum(orc, dist, q):
    A = orc.gen_matrix(20000, 4, dist=dist, seed=3, stream=1)
    B = orc.gen_matrix(20000, 5, dist=dist, seed=3, stream=2)
    assert np.array_equal(orc.atb_exact(A, B, q, q), _fsum_atb(A, B))


@pytest.mark.parametrize("n,k,m,dist", [(100003, 8, 8, 0), (65536, 4, 32, 0), (50000, 3, 5, 1),
                                        (1 << 20, 1, 1, 0)])
def test_dot2_within_one_ulp_of_exact(orc, n, k, m, dist):
    A = orc.gen_matrix(n, k, dist=dist, seed=5, stream=1)
    B = orc.gen_matrix(n, m, dist=dist, seed=5, stream=2)
    q = 52 if dist == 0 else 20
    ex = orc.atb_exact(A, B, q, q)
    d2 = orc.atb_dot2(A, B)
    assert np.all(np.abs(d2 - ex) <= np.spacing(np.abs(ex)))
    # and numpy's GEMM (another summation order) is within its error bound
    _, absC = orc.atb_dot2(A, B, with_abs=True)
    assert np.all(np.abs(A.T @ B - ex) <= n * 2.0**-53 * absC)


def test_exact_oracle_rejects_off_grid(orc):
    A = np.array([[0.1], [0.2]])
    with pytest.raises(ValueError):
        orc.atb_exact(A, A, 20, 20)


def test_naive_is_the_left_to_right_loop(orc):
    A = orc.gen_matrix(300, 3, seed=2, stream=1)
    B = orc.gen_matrix(300, 4, seed=2, stream=2)
    C = orc.atb_naive(A, B)
    for i in range(3):
        for j in range(4):
            acc = 0.0
            for r in range(300):
                acc += float(A[r, i]) * float(B[r, j])
            assert C[i, j] == acc


def test_acceptance_rule(orc):
    ref = np.array([1.0, 1e-12, 100.0, -5.0])
    # diff must exceed BOTH abs_tol (1e-10) and rel_tol*mag (1e-6) to fail
    assert orc.first_mismatch(ref, ref + np.array([5e-7, 5e-11, 5e-5, 0.0])) == -1
    assert orc.first_mismatch(ref, ref + np.array([0.0, 2e-10, 0.0, 0.0])) == 1
    assert orc.first_mismatch(ref, ref + np.array([0.0, 0.0, 2e-4, 0.0])) == 2
    assert orc.first_mismatch(ref, np.array([1.0, 1e-12, 100.0, np.nan])) == 3
    # the generated harness uses abs 1e-4, rel 1e-6 (sycl_codegen_new.py:403-404)
    assert orc.first_mismatch(ref, ref + 5e-5, rel_tol=1e-6, abs_tol=1e-4) == -1


def test_storage_rounding(orc):
    x = orc.gen_matrix(1000, 4, seed=9, stream=1)
    assert np.array_equal(orc.round_f32(x), x.astype(np.float32))
    kat = [(1.0, 0x3F80), (-1.0, 0xBF80), (0.0, 0x0000), (-0.0, 0x8000),
           (1.0 + 2.0**-8, 0x3F80),             # tie -> even (down)
           (1.0 + 3 * 2.0**-8, 0x3F82),         # tie -> even (up)
           (1.0 + 2.0**-8 + 2.0**-30, 0x3F81),  # above the tie: one rounding goes up
           (0.5 - 2.0**-60, 0x3F00), (2.0**-20, 0x3580)]
    for v, bits in kat:
        assert orc.lib().oracle_round_bf16(v) == bits, (v, hex(bits))
    # double rounding FP64 -> FP32 -> BF16 gives 0x3F80 for the third-from-last case
    f32 = np.float32(1.0 + 2.0**-8 + 2.0**-30)
    assert f32 == np.float32(1.0 + 2.0**-8)
    h = orc.round_bf16_bits(x)
    back = orc.bf16_bits_to_f64(h)
    assert np.all(np.abs(back - x) <= 2.0**-8 * np.abs(x))  # half an ulp of 8 bits


def test_backward_closed_form_and_finite_differences(orc):
    n, k, m = 4096, 4, 4
    A = orc.gen_matrix(n, k, seed=13, stream=1)
    B = orc.gen_matrix(n, m, seed=13, stream=2)
    G = orc.gen_matrix(k, m, dist=orc.DIST_NORMAL, seed=13, stream=3)
    dA, dB = orc.atb_backward(A, B, G)
    np.testing.assert_allclose(dA, B @ G.T, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(dB, A @ G, rtol=1e-13, atol=1e-13)

    def L(A_, B_):
        return float(np.sum(G * orc.atb_exact(A_, B_, 60, 60)))

    h = 1e-6  # L is bilinear: central differences are exact up to rounding
    rng = np.random.default_rng(0)
    for r, i in zip(rng.integers(0, n, 6), rng.integers(0, k, 6)):
        Ap, Am = A.copy(), A.copy()
        Ap[r, i] += h
        Am[r, i] -= h
        fd = (L(Ap, B) - L(Am, B)) / (Ap[r, i] - Am[r, i])
        assert abs(fd - dA[r, i]) <= 1e-6 * (1 + abs(dA[r, i]))
        Bp, Bm = B.copy(), B.copy()
        Bp[r, i] += h
        Bm[r, i] -= h
        fd = (L(A, Bp) - L(A, Bm)) / (Bp[r, i] - Bm[r, i])
        assert abs(fd - dB[r, i]) <= 1e-6 * (1 + abs(dB[r, i]))


def test_invariants(orc):
    A = orc.gen_matrix(5000, 6, seed=17, stream=1)
    B = orc.gen_matrix(5000, 3, seed=17, stream=2)
    C = orc.atb_exact(A, B, 52, 52)
    assert np.array_equal(orc.atb_exact(B, A, 52, 52), C.T)  # C(B, A) = C(A, B)^T
    assert not np.array_equal(orc.atb_exact(A[:, :3], B, 52, 52), C[:3].T)  # operands matter
    assert np.array_equal(orc.atb_exact(4 * A, B, 52, 52), 4 * C)  # 2^j scaling is exact
    # split-n additivity: C1 + C2 carries the roundings of C1, C2 and the sum
    C1 = orc.atb_exact(A[:1234], B[:1234], 52, 52)
    C2 = orc.atb_exact(A[1234:], B[1234:], 52, 52)
    bound = np.spacing(np.abs(C1)) + np.spacing(np.abs(C2)) + np.spacing(np.abs(C))
    assert np.all(np.abs(C1 + C2 - C) <= bound)
    # linearity in B
    B2 = orc.gen_matrix(5000, 3, seed=17, stream=3)
    lhs = orc.atb_exact(A, B + B2, 52, 52)
    assert np.all(np.abs(lhs - (C + orc.atb_exact(A, B2, 52, 52))) <= 4 * np.spacing(np.abs(lhs) + 1))
    # special cases: k = m = 1 is a dot product; an all-ones column gives column sums
    ones = np.ones((5000, 1))
    assert np.array_equal(orc.atb_exact(ones, B, 0, 52)[0], [math.fsum(B[:, j]) for j in range(3)])
    G = orc.atb_exact(A, A, 52, 52)
    assert np.array_equal(G, G.T) and np.all(np.linalg.eigvalsh(G) > -1e-9)


def test_hand_checkable_triple(orc):
    A = np.array([[1.0, 2.0], [3.0, 4.0]])
    B = np.array([[5.0], [6.0]])
    expect = np.array([[23.0], [34.0]])
    assert np.array_equal(orc.atb_exact(A, B, 0, 0), expect)
    assert np.array_equal(orc.atb_dot2(A, B), expect)
    assert np.array_equal(orc.atb_naive(A, B), expect)


def test_oracle_reproduces_golden_fixtures(orc):
    """The committed fixtures (tests/golden/make_golden.py) still come out of the oracle."""
    import json
    import os
    import sys

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_golden

    cases = json.load(open(os.path.join(here, "atb_golden.json")))["cases"]
    assert len(cases) == len(make_golden.CASES)
    for case in cases:
        if case["n"] > 70000:
            continue  # the larger cases are exercised by the GPU parity tests
        A, B = make_golden.inputs(case["storage"], case["dist"], case["n"], case["k"], case["m"])
        q = 52 if case["dist"] == 0 else 20
        C = orc.atb_exact(A, B, q, q).ravel()
        assert [float(x).hex() for x in C] == case["C"]


def test_normal_generator_moments(orc):
    """The Box-Muller normal(0, 1) draws (used for the upstream gradient dC):
    mean, variance, skewness and kurtosis within 6 standard errors over 2^20
    draws, and a Kolmogorov-Smirnov test against the normal CDF."""
    from scipy import stats

    x = orc.gen_matrix(1 << 20, 1, dist=orc.DIST_NORMAL, seed=orc.SEED, stream=77).ravel()
    N = x.size
    assert abs(x.mean()) <= 6 / np.sqrt(N)
    assert abs(x.var() - 1) <= 6 * np.sqrt(2 / N)
    assert abs(stats.skew(x)) <= 6 * np.sqrt(6 / N)
    assert abs(stats.kurtosis(x)) <= 6 * np.sqrt(24 / N)  # excess kurtosis 0
    assert stats.kstest(x, "norm").pvalue > 1e-4
    # a planted mistake (the cosine's argument without 2 pi) fails the same checks
    u1 = (np.arange(N) + 0.5) / N
    bad = np.sqrt(-2 * np.log(np.random.default_rng(0).random(N) + 1e-300)) * np.cos(u1)
    assert stats.kstest(bad, "norm").pvalue < 1e-4


def test_bf16_rounding_of_subnormals(orc):
    """Below 2^-126 bf16 values are multiples of 2^-133: known answers of the
    one rounding (to the nearest multiple, ties to even), where rounding to
    FP32 first and truncating would give the neighbour below."""
    u = 2.0**-133
    cases = {-1.5e-40: 0x8002, 1e-39: 0x000B, 3e-39: 0x0021, 2.5 * u: 0x0002, 3.5 * u: 0x0004,
             0.49 * u: 0x0000, 127.5 * u: 0x0080, 2.0**-126: 0x0080, -0.5 * u: 0x8000}
    got = orc.round_bf16_bits(np.array(list(cases)))
    assert [int(g) for g in got] == list(cases.values())
