"""GPU parity of libtsm against the CPU oracle (run on an MI355X: -m gpu).

Tolerances (elementwise, u = unit roundoff of the accumulation type):
  FP64 inputs:       |C - C_exact| <= 4 sqrt(n) 2^-52 (|A|^T |B|)_ij
  FP32/BF16 inputs:  the same form with 2^-23, and the Frobenius relative
                     error <= 1e-6 against the FP64 oracle on the SAME rounded
                     inputs.
Deterministic reduction: the same call gives the same bits on every run.
"""
import ctypes as ct
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import _lib, ops  # noqa: E402

SEED = oracle.SEED
HERE = os.path.dirname(os.path.abspath(__file__))
DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}
EPS = {torch.float64: 2.0**-52, torch.float32: 2.0**-23, torch.bfloat16: 2.0**-23}


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    torch.cuda.set_device(0)
    _lib.lib()
    yield


def host(X):
    """Device matrix -> FP64 numpy (exact for every supported type)."""
    if X.dtype == torch.bfloat16:
        return oracle.bf16_bits_to_f64(X.cpu().view(torch.int16).numpy().view(np.uint16))
    return X.cpu().double().numpy()


def rand(n, w, dt, stream, dist=0):
    return ops.random_matrix(n, w, dt, seed=SEED, stream_id=stream, dist=dist)


def check_close(C, Ah, Bh, dt, exact=None):
    """C (k x m numpy) against the oracle on the host copies of the inputs."""
    n = Ah.shape[0]
    ref, absC = oracle.atb_dot2(Ah, Bh, with_abs=True)
    if exact is not None:
        ref = exact
    bound = 4.0 * np.sqrt(max(n, 1)) * EPS[dt] * absC
    err = np.abs(C - ref)
    assert np.all(err <= bound), f"max err/bound {np.max(err / np.maximum(bound, 1e-300)):.3g}"
    if dt != torch.float64 and np.linalg.norm(ref) > 0:
        assert np.linalg.norm(C - ref) / np.linalg.norm(ref) <= 1e-6
    return err


# ---- inputs ------------------------------------------------------------------------
@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("dist", [0, 1])
def test_fill_is_bit_exact(dt, dist):
    X = rand(3001, 7, DT[dt], 3, dist)
    ref = oracle.gen_matrix(3001, 7, dist=dist, seed=SEED, stream=3)
    if dt == "f64":
        assert np.array_equal(X.cpu().numpy().view(np.uint64), ref.view(np.uint64))
    elif dt == "f32":
        assert np.array_equal(X.cpu().numpy().view(np.uint32), oracle.round_f32(ref).view(np.uint32))
    else:
        got = X.cpu().view(torch.int16).numpy().view(np.uint16)
        assert np.array_equal(got, oracle.round_bf16_bits(ref))


# ---- golden fixtures ----------------------------------------------------------------
GOLDEN = json.load(open(os.path.join(HERE, "golden", "atb_golden.json")))["cases"]


@pytest.mark.parametrize("case", GOLDEN, ids=lambda c: f"{c['storage']}-d{c['dist']}-n{c['n']}-{c['k']}x{c['m']}")
def test_golden_fixtures(case):
    dt = DT[case["storage"]]
    n, k, m = case["n"], case["k"], case["m"]
    A = rand(n, k, dt, 1, case["dist"])
    B = rand(n, m, dt, 2, case["dist"])
    C = ops.atb(A, B, out_dtype=torch.float64).cpu().numpy().ravel()
    exact = np.array([float.fromhex(x) for x in case["C"]])
    absC = np.array([float.fromhex(x) for x in case["absC"]])
    bound = 4.0 * np.sqrt(n) * EPS[dt] * absC
    assert np.all(np.abs(C - exact) <= bound)
    if dt != torch.float64:
        assert np.linalg.norm(C - exact) / np.linalg.norm(exact) <= 1e-6


# ---- shapes and layouts -------------------------------------------------------------------
# every final-reduction regime: 1 output, 2..256 (one pair per thread or fewer),
# 512 (two outputs per thread), > 512 (per-output sums)
SHAPES = [(1, 1), (1, 2), (2, 2), (3, 5), (8, 8), (4, 32), (32, 4), (16, 16), (16, 32), (32, 16),
          (8, 64), (33, 17), (64, 64), (1, 64), (64, 1), (7, 9)]
NS = [0, 1, 2, 255, 256, 257, 4097, 100003]


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("km", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_shapes_and_ragged_rows(dt, km):
    k, m = km
    for n in NS:
        A, B = rand(n, k, DT[dt], 1), rand(n, m, DT[dt], 2)
        C = ops.atb(A, B, out_dtype=torch.float64).cpu().numpy()
        if n == 0:
            assert np.array_equal(C, np.zeros((k, m)))
            continue
        check_close(C, host(A), host(B), DT[dt])


@pytest.mark.parametrize("il", [0, 1])
@pytest.mark.parametrize("grid", [1, 3, 1000])
def test_tile_orders_and_grids(il, grid):
    for n in (1, 255, 256, 257, 5000, 65536 + 300):
        A, B = rand(n, 8, torch.float64, 1), rand(n, 8, torch.float64, 2)
        C = ops.atb(A, B, config=_lib.make_config(grid=grid, interleave=il)).cpu().numpy()
        check_close(C, host(A), host(B), torch.float64)
        S = A[:, :5]  # strided (element-load path) with the same tile orders
        C = ops.atb(S, B, config=_lib.make_config(grid=grid, interleave=il)).cpu().numpy()
        check_close(C, host(S), host(B), torch.float64)


@pytest.mark.parametrize("dt,k,m", [("f64", 8, 8), ("f64", 2, 2), ("f32", 16, 16), ("bf16", 8, 4)])
def test_final_sum_protocols_at_their_boundary(dt, k, m):
    """Grids below 256 workgroups (and below 16384 partial words) are summed by
    the streaming kernel's last workgroup; from 256 on (or 16384 words) by the
    finisher kernel.  Both sides of each boundary against the oracle, plain and
    accumulating; a grid forced through the config reaches the same paths."""
    kcmc = (1 << (k - 1).bit_length()) * (1 << (m - 1).bit_length())
    idx, _ = _lib.plan(_lib.F64 if dt == "f64" else _lib.F32 if dt == "f32" else _lib.BF16, 1, k, m)
    rows = _lib.kernels()[idx]["rows"]
    small = min(255, 16383 // kcmc)
    for grid in (small, small + 1, 256, 1000):
        n = grid * rows - 7  # the last tile is partial
        A, B = rand(n, k, DT[dt], 11), rand(n, m, DT[dt], 12)
        cfg = _lib.make_config(grid=grid)
        C = ops.atb(A, B, out_dtype=torch.float64, config=cfg).cpu().numpy()
        check_close(C, host(A), host(B), DT[dt])
        C0 = torch.full((k, m), 0.25, dtype=torch.float64, device="cuda")
        ops.atb(A, B, out=C0, accumulate=True, config=cfg)
        ref = C + 0.25
        assert np.all(np.abs(C0.cpu().numpy() - ref) <= np.spacing(np.abs(ref)) * 2)


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
def test_strided_views_and_misaligned_bases(dt):
    n = 50021
    X = rand(n + 1, 24, DT[dt], 5)
    Y = rand(n + 1, 16, DT[dt], 6)
    cases = [
        (X[:n, 3:11], Y[:n, 0:8]),   # lda = 24 > k, ldb = 16 > m
        (X[1:, :8], Y[1:, :16]),     # base offset by one row, lda = 24 > k
        (X[:n, 1:9], Y[:n, 1:5]),    # base offset by one element (misaligned for vector loads)
    ]
    for A, B in cases:
        C = ops.atb(A, B, out_dtype=torch.float64).cpu().numpy()
        check_close(C, host(A), host(B), DT[dt])


def test_accumulate_and_f32_output():
    n, k, m = 70001, 8, 8
    A, B = rand(n, k, torch.float64, 1), rand(n, m, torch.float64, 2)
    ref, absC = oracle.atb_dot2(host(A), host(B), with_abs=True)
    C0 = torch.from_numpy(oracle.gen_matrix(k, m, seed=SEED, stream=9)).cuda()
    C = C0.clone()
    ops.atb(A, B, out=C, accumulate=True)
    got = C.cpu().numpy()
    bound = 4 * np.sqrt(n) * 2.0**-52 * absC + np.spacing(np.abs(ref + C0.cpu().numpy()))
    assert np.all(np.abs(got - (ref + C0.cpu().numpy())) <= bound)
    C32 = ops.atb(A, B, out_dtype=torch.float32).cpu().numpy().astype(np.float64)
    assert np.all(np.abs(C32 - ref) <= np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
                  + 4 * np.sqrt(n) * 2.0**-52 * absC)
    # accumulate into an FP32 output
    D = torch.ones((k, m), dtype=torch.float32, device="cuda")
    ops.atb(A, B, out=D, accumulate=True)
    assert np.all(np.abs(D.cpu().numpy() - (ref + 1.0)) <= 2 * np.spacing(np.float32(np.abs(ref).max() + 1)))


@pytest.mark.parametrize("dt,k,m", [("f64", 8, 8), ("f64", 4, 32), ("f64", 64, 64),
                                    ("f64", 32, 32), ("f64", 16, 16), ("f32", 32, 32),
                                    ("f32", 64, 64), ("bf16", 32, 32), ("bf16", 64, 64),
                                    ("bf16", 16, 16), ("f64", 1, 1)])
def test_every_generated_variant(dt, k, m):
    dtype_id = {"f64": 0, "f32": 1, "bf16": 2}[dt]
    n = 123457
    A, B = rand(n, k, DT[dt], 1), rand(n, m, DT[dt], 2)
    Ah, Bh = host(A), host(B)
    seen = 0
    for kd in _lib.kernels():
        if kd["in_type"] != dtype_id or kd["kc"] != k or kd["mc"] != m:
            continue
        for grid in (0, 7):
            for il in (0, 1):
                cfg = _lib.config_of(kd, grid=grid, interleave=il)
                C = ops.atb(A, B, out_dtype=torch.float64, config=cfg).cpu().numpy()
                check_close(C, Ah, Bh, DT[dt])
        seen += 1
    assert seen >= 1


# classes with a side narrower than 16 whose rows are whole 16-byte chunks
# (or, BF16 at 4, 8 bytes) beside a side of 16 or more
PADDED_CLASSES = [(dt, k, m) for dt, narrow in (("f64", (2, 4, 8)), ("f32", (4, 8)), ("bf16", (4, 8)))
                  for w in narrow for x in (16, 32, 64) for k, m in ((w, x), (x, w))]


@pytest.mark.parametrize("dt,k,m", PADDED_CLASSES, ids=lambda x: str(x))
def test_padded_matrix_core_kernels(dt, k, m):
    """The matrix-core kernels of the classes with one narrow side (that
    side's rows staged into LDS rows of 16, the rest zero), forced one
    by one: dense inputs on several grids and tile orders, strided views
    with fewer columns than the class (the element path), and sign * 2^e
    inputs against the exact int128 oracle."""
    dtype_id = {"f64": 0, "f32": 1, "bf16": 2}[dt]
    kds = [kd for kd in _lib.kernels() if kd["in_type"] == dtype_id and kd["kc"] == k
           and kd["mc"] == m and kd["engine"] == _lib.ENGINE_MFMA_LDS]
    assert kds, "no padded matrix-core kernel compiled for this class"
    n = 70001
    A, B = rand(n, k, DT[dt], 1), rand(n, m, DT[dt], 2)
    Ah, Bh = host(A), host(B)
    # views on misaligned bases with one column fewer than the class where
    # that stays in the class (else all of it, ld > width): the element path
    kw, mw = (k - 1 if k > 2 else k), (m - 1 if m > 2 else m)
    Aw = rand(5003, k + 3, DT[dt], 3)[:, 1:1 + kw]
    Bw = rand(5003, m + 5, DT[dt], 4)[:, 2:2 + mw]
    P, Q = rand(20011, k, DT[dt], 5, dist=1), rand(20011, m, DT[dt], 6, dist=1)
    Ph, Qh = host(P), host(Q)
    exact = oracle.atb_exact(Ph, Qh, 20, 20)
    for kd in kds:
        for grid, il in ((0, 0), (0, 1), (7, 1), (300, 0)):
            cfg = _lib.config_of(kd, grid=grid, interleave=il)
            check_close(ops.atb(A, B, out_dtype=torch.float64, config=cfg).cpu().numpy(),
                        Ah, Bh, DT[dt])
        cfg = _lib.config_of(kd)
        check_close(ops.atb(Aw, Bw, out_dtype=torch.float64, config=cfg).cpu().numpy(),
                    host(Aw), host(Bw), DT[dt])
        check_close(ops.atb(P, Q, out_dtype=torch.float64, config=cfg).cpu().numpy(),
                    Ph, Qh, DT[dt], exact=exact)


def test_unknown_variant_is_reported():
    A, B = rand(100, 8, torch.float64, 1), rand(100, 8, torch.float64, 2)
    with pytest.raises(_lib.TsmError) as e:
        ops.atb(A, B, config=_lib.make_config(tm=8, tn=8, rows=12345))
    assert e.value.status == _lib.ERR_NO_KERNEL


def test_determinism_bitwise():
    A, B = rand(1 << 20, 8, torch.float64, 1), rand(1 << 20, 8, torch.float64, 2)
    first = ops.atb(A, B)
    s = torch.cuda.Stream()
    for i in range(20):
        with torch.cuda.stream(s if i % 2 else torch.cuda.current_stream()):
            C = ops.atb(A, B)
        torch.cuda.synchronize()
        assert torch.equal(C, first)


@pytest.mark.parametrize("dt,k,m,n", [
    ("f64", 8, 8, 1 << 22),    # register tiles, finisher kernel (the headline shape)
    ("f64", 1, 1, 1 << 20),    # register tiles, finisher kernel
    ("f64", 8, 8, 20000),      # register tiles, final sum inside the streaming kernel
    ("f64", 64, 64, 1 << 18),  # direct matrix-core engine
    ("f32", 32, 32, 1 << 20),  # LDS-staged matrix-core engine, (hi, lo) folds
    ("bf16", 64, 64, 1 << 20), # BF16 direct engine (wave-private LDS tiles)
])
def test_determinism_bitwise_every_engine(dt, k, m, n):
    """100 calls give the same bits, on alternating streams while a copy
    kernel on a third stream competes for the compute units (so workgroups
    start, finish and arrive in different orders from call to call)."""
    A, B = rand(n, k, DT[dt], 1), rand(n, m, DT[dt], 2)
    first = ops.atb(A, B, out_dtype=torch.float64)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    noise = torch.cuda.Stream()
    src = torch.empty(1 << 24, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    outs = []
    for i in range(100):
        if i % 10 == 0:
            with torch.cuda.stream(noise):
                dst.copy_(src)
        with torch.cuda.stream(streams[i % 2]):
            outs.append(ops.atb(A, B, out_dtype=torch.float64))
    torch.cuda.synchronize()
    assert all(torch.equal(C, first) for C in outs)


def test_determinism_bitwise_backward_and_batched():
    """The gradients and a batched launch give the same bits over 50 calls."""
    n = 300_001
    G = torch.from_numpy(oracle.gen_matrix(16, 16, dist=oracle.DIST_NORMAL, seed=SEED, stream=3))
    for dt in ("f64", "f32", "bf16"):
        A, B = rand(n, 16, DT[dt], 1), rand(n, 16, DT[dt], 2)
        g = G.to("cuda", torch.float64 if dt == "f64" else torch.float32)
        dA0, dB0 = ops.atb_backward(A, B, g)
        for _ in range(50):
            dA, dB = ops.atb_backward(A, B, g)
            assert torch.equal(dA, dA0) and torch.equal(dB, dB0)
    As = [rand(int(n * f), w, torch.float64, 1) for f, w in ((1, 8), (0.3, 32), (2.1, 4), (0.01, 64))]
    Bs = [rand(X.shape[0], w, torch.float64, 2) for X, w in zip(As, (8, 16, 32, 2))]
    first = ops.atb_batched(As, Bs)
    for _ in range(50):
        assert all(torch.equal(C, C0) for C, C0 in zip(ops.atb_batched(As, Bs), first))


def test_size_independent_properties():
    n = 3_000_017
    A, B = rand(n, 8, torch.float64, 1), rand(n, 8, torch.float64, 2)
    C = ops.atb(A, B)
    # transposition: the same kernel with the operands swapped gives C^T bit for bit
    assert torch.equal(ops.atb(B, A), C.T)
    # power-of-two scaling is exact
    assert torch.equal(ops.atb(4 * A, B), 4 * C)
    # split-n additivity through accumulate
    h = 1_234_567
    D = ops.atb(A[:h], B[:h])
    ops.atb(A[h:], B[h:], out=D, accumulate=True)
    _, absC = oracle.atb_dot2(host(A), host(B), with_abs=True)
    assert np.all(np.abs((D - C).cpu().numpy()) <= 8 * np.sqrt(n) * 2.0**-52 * absC)


@pytest.mark.parametrize("n,k,m", [(1 << 25, 8, 8), (14_913_080, 4, 32), (14_913_080, 32, 4)])
def test_full_size_sweep_against_exact_oracle(n, k, m):
    """The FP64 sweep shapes at A + B = 4 GiB against the exact int128 oracle
    (the host regenerates the same inputs from the counter-based generator)."""
    A, B = rand(n, k, torch.float64, 1), rand(n, m, torch.float64, 2)
    C = ops.atb(A, B).cpu().numpy()
    del A, B
    Ah = oracle.gen_matrix(n, k, seed=SEED, stream=1)
    Bh = oracle.gen_matrix(n, m, seed=SEED, stream=2)
    exact = oracle.atb_exact(Ah, Bh, 52, 52)
    check_close(C, Ah, Bh, torch.float64, exact=exact)


def test_row_index_beyond_2_to_31_elements():
    """n * k > 2^31: 64-bit addressing, checked by split-n additivity over byte
    offsets past 4 GiB (FP32 storage, 8 GiB per matrix)."""
    n, k = (1 << 28) + 123, 8
    A = rand(n, k, torch.float32, 1)
    B = rand(n, k, torch.float32, 2)
    full = ops.atb(A, B, out_dtype=torch.float64)
    parts = sum(ops.atb(A[i:j], B[i:j], out_dtype=torch.float64)
                for i, j in ((0, 1 << 26), (1 << 26, 3 << 26), (3 << 26, n)))
    tail = ops.atb(A[n - 1000:], B[n - 1000:], out_dtype=torch.float64).cpu().numpy()
    check_close(tail, host(A[n - 1000:]), host(B[n - 1000:]), torch.float32)
    # Both sums carry the FP32 block-accumulation error (about sqrt(32) * 2^-24
    # relative for 32-row blocks), so they agree to the FP32 target, 1e-6.
    rel = (torch.linalg.norm(full - parts) / torch.linalg.norm(full)).item()
    assert rel <= 1e-6


# ---- precision (FP32 / BF16 storage) ---------------------------------------------------------
@pytest.mark.parametrize("dt", ["f32", "bf16"])
@pytest.mark.parametrize("dist", [0, 1])
def test_low_precision_storage_error(dt, dist):
    n, k, m = 1 << 22, 32, 32
    A, B = rand(n, k, DT[dt], 1, dist), rand(n, m, DT[dt], 2, dist)
    C = ops.atb(A, B, out_dtype=torch.float64).cpu().numpy()
    Ah, Bh = host(A), host(B)
    q = 52 if dist == 0 else 20
    exact = oracle.atb_exact(Ah, Bh, q, q)
    _, absC = oracle.atb_dot2(Ah, Bh, with_abs=True)
    frob = np.linalg.norm(C - exact) / np.linalg.norm(exact)
    assert frob <= 1e-6, frob
    assert np.max(np.abs(C - exact) / absC) <= 1e-6


# ---- backward ------------------------------------------------------------------------------------
@pytest.mark.parametrize("dt,k,m", [("f64", 4, 4), ("f64", 16, 16), ("f64", 64, 3), ("f64", 5, 64),
                                    ("f32", 16, 16), ("bf16", 8, 8), ("f32", 1, 1)])
def test_backward_against_oracle(dt, k, m):
    n = 40001
    A, B = rand(n, k, DT[dt], 1), rand(n, m, DT[dt], 2)
    G = torch.from_numpy(oracle.gen_matrix(k, m, dist=oracle.DIST_NORMAL, seed=SEED, stream=3))
    gdt = torch.float64 if dt == "f64" else torch.float32
    dA, dB = ops.atb_backward(A, B, G.to("cuda", gdt))
    dA_ref, dB_ref = oracle.atb_backward(host(A), host(B), G.to(gdt).double().numpy())
    if dt == "f64":
        tol = dict(rtol=1e-13, atol=1e-13)
    elif dt == "f32":
        tol = dict(rtol=1e-5, atol=1e-5)
    else:
        tol = dict(rtol=2**-7, atol=2**-7 * np.abs(G.numpy()).sum())
    np.testing.assert_allclose(host(dA), dA_ref, **tol)
    np.testing.assert_allclose(host(dB), dB_ref, **tol)
    # one side only
    dA2, dB2 = ops.atb_backward(A, B, G.to("cuda", gdt), need_dB=False)
    assert dB2 is None and torch.equal(dA2, dA)


def valu_class(dt, k, m):
    """Dense rows of this shape take a VALU kernel (staged or packed, widths
    up to 8), whose bits equal the row-wise kernel's."""
    return max(k, m) <= 8


@pytest.mark.parametrize("dt,k,m,n", [("f32", 16, 16, 70001), ("f32", 12, 4, 3001), ("f32", 32, 32, 300),
                                      ("f64", 2, 16, 5), ("f64", 16, 16, 4097), ("bf16", 8, 8, 1000),
                                      ("bf16", 64, 16, 2049), ("bf16", 4, 4, 100003), ("bf16", 4, 4, 3),
                                      ("bf16", 2, 2, 1001), ("bf16", 1, 1, 37), ("f32", 2, 2, 70001),
                                      ("f32", 1, 1, 5), ("f64", 1, 1, 3001),
                                      # 32-byte rows: two rows per lane in the staged kernel
                                      ("f64", 4, 4, 70001), ("f64", 4, 2, 129), ("f32", 8, 8, 100003),
                                      ("f32", 8, 4, 127), ("bf16", 16, 16, 40001), ("bf16", 16, 8, 257),
                                      ("bf16", 8, 16, 64),
                                      # one 16-byte row per chunk in the packed kernel
                                      ("f64", 2, 2, 70001), ("f64", 2, 2, 3), ("f32", 4, 4, 5001)])
def test_backward_staged_equals_rowwise(dt, k, m, n):
    """Dense rows take the LDS-staged kernel (widths up to 8), the packed
    kernel (square classes with rows under 16 bytes) or a matrix-core kernel
    (16, 32 and 64); strided views take the row-wise one.  The staged, packed and row-wise kernels add the same
    products in the same order, so their bits agree; the matrix-core kernels
    are held to the oracle tolerance."""
    A, B = rand(n, k, DT[dt], 1), rand(n, m, DT[dt], 2)
    G = torch.from_numpy(oracle.gen_matrix(k, m, dist=oracle.DIST_NORMAL, seed=SEED, stream=3))
    gdt = torch.float64 if dt == "f64" else torch.float32
    dA, dB = ops.atb_backward(A, B, G.to("cuda", gdt))
    Aw = torch.zeros((n, k + 3), dtype=A.dtype, device="cuda")
    Bw = torch.zeros((n, m + 5), dtype=B.dtype, device="cuda")
    Aw[:, :k] = A
    Bw[:, :m] = B
    dA2, dB2 = ops.atb_backward(Aw[:, :k], Bw[:, :m], G.to("cuda", gdt))
    if valu_class(dt, k, m):
        assert torch.equal(dA, dA2) and torch.equal(dB, dB2)
    dA_ref, dB_ref = oracle.atb_backward(host(A), host(B), G.to(gdt).double().numpy())
    tol = 1e-13 if dt == "f64" else (1e-5 if dt == "f32" else 2**-7 * np.abs(G.numpy()).sum())
    np.testing.assert_allclose(host(dA), dA_ref, rtol=tol if dt != "bf16" else 2**-7, atol=tol)
    np.testing.assert_allclose(host(dB), dB_ref, rtol=tol if dt != "bf16" else 2**-7, atol=tol)


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("k,m,n", [(32, 32, 40001), (64, 64, 20011), (40, 24, 1000), (64, 8, 129),
                                   (8, 64, 127), (48, 48, 5), (32, 64, 128), (32, 32, 63),
                                   (24, 32, 257), (32, 16, 5), (32, 32, 513), (16, 16, 257),
                                   (16, 4, 1000), (12, 16, 129), (16, 8, 255)])
def test_backward_matrix_core_against_oracle(dt, k, m, n):
    """The 32- and 64-wide classes, and the FP64/FP32 16-wide ones (dense
    rows), take a matrix-core kernel (tiles of 64 rows for FP64 W = 32, 256
    for BF16 W = 32 and FP32 W = 16, 128 otherwise: the row counts cover one
    partial tile, a tile plus one row and several tiles plus a partial one)."""
    if (k * DT[dt].itemsize) % 16 or (m * DT[dt].itemsize) % 16:
        pytest.skip("rows that are not whole 16-byte chunks take the row-wise kernel")
    A, B = rand(n, k, DT[dt], 1), rand(n, m, DT[dt], 2)
    G = torch.from_numpy(oracle.gen_matrix(k, m, dist=oracle.DIST_NORMAL, seed=SEED, stream=3))
    gdt = torch.float64 if dt == "f64" else torch.float32
    dA, dB = ops.atb_backward(A, B, G.to("cuda", gdt))
    dA_ref, dB_ref = oracle.atb_backward(host(A), host(B), G.to(gdt).double().numpy())
    if dt == "f64":
        tol = dict(rtol=1e-13, atol=1e-13)
    elif dt == "f32":
        tol = dict(rtol=1e-5, atol=1e-5)
    else:
        tol = dict(rtol=2**-7, atol=2**-7 * np.abs(G.numpy()).sum())
    np.testing.assert_allclose(host(dA), dA_ref, **tol)
    np.testing.assert_allclose(host(dB), dB_ref, **tol)
    _, dB1 = ops.atb_backward(A, B, G.to("cuda", gdt), need_dA=False)
    assert torch.equal(dB1, dB)


@pytest.mark.parametrize("k,m,n", [(64, 64, 20011), (32, 32, 40001), (40, 24, 1000), (64, 8, 129),
                                   (8, 64, 127), (48, 48, 5), (32, 32, 255), (24, 32, 257),
                                   (32, 16, 5), (16, 16, 40001), (16, 8, 513), (8, 16, 511)])
def test_backward_bf16_matrix_core_against_fp32_path(k, m, n):
    """The BF16 matrix-core kernel (dC split exactly into three BF16 pieces,
    FP32 accumulation in the matrix instruction) against the row-wise kernel
    on a strided view (FP32 products added in column order).  Both FP32 sums
    are within 64 * 2^-24 * S of the exact one (S = sum of |products|), so
    the BF16 outputs differ by at most one BF16 rounding (2^-7 relative) plus
    2^-17 S, and almost all are equal.  Dropping the middle piece would move
    outputs by about 2^-8 S."""
    A, B = rand(n, k, torch.bfloat16, 1), rand(n, m, torch.bfloat16, 2)
    G = torch.from_numpy(oracle.gen_matrix(k, m, dist=oracle.DIST_NORMAL, seed=SEED, stream=3))
    Gd = G.to("cuda", torch.float32)
    dA, dB = ops.atb_backward(A, B, Gd)
    Aw = torch.zeros((n, k + 3), dtype=A.dtype, device="cuda")
    Bw = torch.zeros((n, m + 5), dtype=B.dtype, device="cuda")
    Aw[:, :k] = A
    Bw[:, :m] = B
    dA2, dB2 = ops.atb_backward(Aw[:, :k], Bw[:, :m], Gd)
    Gabs = np.abs(Gd.cpu().double().numpy())
    S_A = np.abs(host(B)) @ Gabs.T   # dA = B dC^T
    S_B = np.abs(host(A)) @ Gabs     # dB = A dC
    for name, X, Y, S in (("dA", dA, dA2, S_A), ("dB", dB, dB2, S_B)):
        x, y = host(X), host(Y)
        err = np.abs(x - y)
        tol = 2.0**-7 * np.maximum(np.abs(x), np.abs(y)) + 2.0**-17 * S
        worst = np.unravel_index(np.argmax(err / np.maximum(tol, 1e-300)), err.shape)
        assert np.all(err <= tol), (name, worst, x[worst], y[worst], S[worst])
        same = np.mean(x == y)
        assert same >= 0.995, (name, same)


@pytest.mark.parametrize("dt,k,m", [("f64", 8, 8), ("f64", 32, 32), ("f64", 64, 64), ("f32", 16, 48),
                                    ("bf16", 64, 64)])
def test_torch_front_end_in_a_captured_graph(dt, k, m):
    """ops.atb and ops.ata with a caller workspace (ops.workspace) captured in a
    torch CUDA graph: every replay, with eager calls between them, gives the
    eager bits (in-kernel final sum, finisher kernel and direct engine)."""
    n = 300007
    A, B = rand(n, k, DT[dt], 1), rand(n, m, DT[dt], 2)
    ws = ops.workspace(n, k, m, DT[dt])
    ws2 = ops.workspace(n, k, None, DT[dt])
    C = torch.empty((k, m), dtype=torch.float64, device="cuda")
    S = torch.empty((k, k), dtype=torch.float64, device="cuda")
    ops.atb(A, B, out=C, workspace=ws)
    ops.ata(A, out=S, workspace=ws2)
    torch.cuda.synchronize()
    eager_c, eager_s = C.clone(), S.clone()
    assert torch.equal(eager_c, ops.atb(A, B, out_dtype=torch.float64))  # the pool gives the same bits
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ops.atb(A, B, out=C, workspace=ws)
        ops.ata(A, out=S, workspace=ws2)
    for i in range(4):
        C.zero_()
        S.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, eager_c) and torch.equal(S, eager_s), f"replay {i}"
        ops.atb(A, B, out=C, workspace=ws)  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(C, eager_c)


def test_autograd_matches_torch_matmul():
    n, k, m = 30011, 16, 16
    A0, B0 = rand(n, k, torch.float64, 1), rand(n, m, torch.float64, 2)
    G = torch.from_numpy(oracle.gen_matrix(k, m, dist=oracle.DIST_NORMAL, seed=SEED, stream=3)).cuda()
    A, B = A0.clone().requires_grad_(), B0.clone().requires_grad_()
    (ops.atb_autograd(A, B) * G).sum().backward()
    At, Bt = A0.clone().requires_grad_(), B0.clone().requires_grad_()
    (torch.matmul(At.T.contiguous(), Bt) * G).sum().backward()
    torch.testing.assert_close(A.grad, At.grad, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(B.grad, Bt.grad, rtol=1e-12, atol=1e-12)
    # the registered operator differentiates the same way
    A2, B2 = A0.clone().requires_grad_(), B0.clone().requires_grad_()
    (torch.ops.tsm.atb(A2, B2) * G).sum().backward()
    assert torch.equal(A2.grad, A.grad) and torch.equal(B2.grad, B.grad)


def test_gradcheck_fp64():
    A = rand(64, 3, torch.float64, 1).requires_grad_()
    B = rand(64, 4, torch.float64, 2).requires_grad_()
    assert torch.autograd.gradcheck(lambda a, b: ops.atb_autograd(a, b), (A, B), eps=1e-6,
                                    atol=1e-8)


# ---- graph capture -------------------------------------------------------------------------------
@pytest.mark.parametrize("k,m", [(8, 8), (32, 32), (64, 64)])
def test_hip_graph_replays_match_eager(k, m):
    """In-kernel final sum (8 x 8), the finisher kernel after the staged
    matrix-core kernel (32 x 32) and after the direct engine (64 x 64)."""
    n = 1 << 20
    A, B = rand(n, k, torch.float64, 1), rand(n, m, torch.float64, 2)
    L = _lib.lib()
    ws_bytes = L.tsm_atb_workspace_size(_lib.F64, n, k, m, None)
    ws = torch.zeros(ws_bytes, dtype=torch.uint8, device="cuda")
    C = torch.empty((k, m), dtype=torch.float64, device="cuda")

    def call():
        st = L.tsm_atb(_lib.F64, _lib.F64, n, k, m, A.data_ptr(), k, B.data_ptr(), m,
                       C.data_ptr(), m, 0, ws.data_ptr(), ws_bytes,
                       torch.cuda.current_stream().cuda_stream)
        assert st == _lib.SUCCESS

    call()
    torch.cuda.synchronize()
    eager = C.clone()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        call()  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        call()
    for i in range(5):
        C.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, eager), f"replay {i}"
        call()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(C, eager)
