"""GPU parity of the two sibling products (run on an MI355X: -m gpu).

  tsm_ata   C = A^T A from half the bytes of tsm_atb(A, A): checked against
            the oracle with the A^T B tolerances, against tsm_atb(A, A)
            within twice that bound, and for exact symmetry.
  tsm_tsmm  Y = A X: the dB = A dC half of the backward kernels; checked
            against the oracle's closed form (A @ X in FP64) and against
            tsm_atb_backward's dB bit for bit.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import _lib, ops  # noqa: E402

SEED = oracle.SEED
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
    if X.dtype == torch.bfloat16:
        return oracle.bf16_bits_to_f64(X.cpu().view(torch.int16).numpy().view(np.uint16))
    return X.cpu().double().numpy()


def rand(n, w, dt, stream, dist=0):
    return ops.random_matrix(n, w, dt, seed=SEED, stream_id=stream, dist=dist)


# ---- A^T A -------------------------------------------------------------------------------
def check_ata(C, A, n, dt):
    Ah = host(A)
    ref, absC = oracle.atb_dot2(Ah, Ah, with_abs=True)
    Ch = C.cpu().numpy()
    bound = 4.0 * np.sqrt(max(n, 1)) * EPS[DT[dt]] * absC
    assert np.all(np.abs(Ch - ref) <= bound)
    if dt != "f64" and np.linalg.norm(ref) > 0:
        assert np.linalg.norm(Ch - ref) / np.linalg.norm(ref) <= 1e-6
    assert np.array_equal(Ch, Ch.T)  # exactly symmetric
    C2 = ops.atb(A, A, out_dtype=torch.float64).cpu().numpy()
    assert np.all(np.abs(Ch - C2) <= 2 * bound)


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("k", [1, 3, 8, 16, 24, 32, 40, 64])
def test_ata_against_oracle_and_atb(dt, k):
    n = 30011
    A = rand(n, k, DT[dt], 1)
    check_ata(ops.ata(A, out_dtype=torch.float64), A, n, dt)


@pytest.mark.parametrize("dt,k", [("f64", 8), ("f32", 12), ("bf16", 32), ("f64", 64)])
def test_ata_strided_view_and_small_n(dt, k):
    """Element loads (lda > k) and row counts below one tile."""
    for n in (1, 5, 257, 4099):
        W = rand(n, k + 3, DT[dt], 4)
        A = W[:, :k]
        check_ata(ops.ata(A, out_dtype=torch.float64), A, n, dt)
        # the dense copy takes the staged path: within the bound of the strided one
        check_ata(ops.ata(A.contiguous(), out_dtype=torch.float64), A.contiguous(), n, dt)


def test_ata_accumulate_and_fp32_output():
    n, k = 100003, 16
    A = rand(n, k, torch.float64, 5)
    C0 = torch.from_numpy(oracle.gen_matrix(k, k, seed=SEED, stream=6)).cuda()
    C = C0.clone()
    ops.ata(A, out=C, accumulate=True)
    assert torch.equal(C, C0 + ops.ata(A))
    C32 = ops.ata(A, out_dtype=torch.float32)
    assert torch.equal(C32, ops.ata(A).float())


def test_ata_pow2_exact():
    """sign * 2^e inputs: the exact int128 oracle, at 4 GiB-class n."""
    n, k = 1 << 24, 16
    A = rand(n, k, torch.float64, 7, dist=1)
    C = ops.ata(A).cpu().numpy()
    Ah = host(A)
    exact = oracle.atb_exact(Ah, Ah, 20, 20)
    _, absC = oracle.atb_dot2(Ah, Ah, with_abs=True)
    assert np.all(np.abs(C - exact) <= 4.0 * np.sqrt(n) * 2.0**-52 * absC)


def test_ata_64bit_addressing():
    """n * k > 2^31 elements: split-n additivity against two halves."""
    n, k = (1 << 28) + 5, 8
    A = rand(n, k, torch.float32, 8)
    C = ops.ata(A, out_dtype=torch.float64)
    h = n // 2
    C1 = ops.ata(A[:h], out_dtype=torch.float64)
    C2 = ops.ata(A[h:], out_dtype=torch.float64)
    ref = C1 + C2
    assert torch.allclose(C, ref, rtol=1e-6, atol=1e-6 * ref.abs().max().item())
    del A


@pytest.mark.parametrize("dt,rows", [("f64", 128), ("f64", 256), ("f32", 256), ("bf16", 512),
                                     ("bf16", 1024), ("bf16", 2048)])
@pytest.mark.parametrize("pf", [1, 2])
def test_ata_padded_matrix_core_kernels(dt, rows, pf):
    """A^T A at k = 8 on the staged matrix-core consumer with A's rows padded
    to 16 in LDS (one diagonal block), forced: dense and ragged inputs on
    several grids, a strided view of fewer columns (the element path) and
    sign * 2^e inputs against the exact oracle; C exactly symmetric."""
    k = 8
    cfg = dict(rows=rows, pf=pf, nt=True, engine=_lib.ENGINE_MFMA_LDS)
    for n in (70001, 1000):
        A = rand(n, k, DT[dt], 1)
        for grid, il in ((0, 0), (0, 1), (7, 1)):
            C = ops.ata(A, out_dtype=torch.float64,
                        config=_lib.make_config(grid=grid, interleave=il, **cfg))
            check_ata(C, A, n, dt)
    W = rand(5003, k + 3, DT[dt], 3)[:, 1:k]  # 7 columns, misaligned base
    check_ata(ops.ata(W, out_dtype=torch.float64, config=_lib.make_config(**cfg)), W, 5003, dt)
    P = rand(20011, k, DT[dt], 5, dist=1)
    Ph = host(P)
    exact = oracle.atb_exact(Ph, Ph, 20, 20)
    absC = oracle.atb_exact(np.abs(Ph), np.abs(Ph), 20, 20)
    C = ops.ata(P, out_dtype=torch.float64, config=_lib.make_config(**cfg)).cpu().numpy()
    assert np.array_equal(C, C.T)
    assert np.all(np.abs(C - exact) <= 4.0 * np.sqrt(20011) * EPS[DT[dt]] * absC)


def test_ata_deterministic():
    A = rand(1 << 22, 8, torch.float64, 9)
    first = ops.ata(A)
    for _ in range(3):
        assert torch.equal(ops.ata(A), first)


# ---- Y = A X -----------------------------------------------------------------------------
@pytest.mark.parametrize("dt,k,m", [("f64", 8, 8), ("f64", 2, 16), ("f64", 64, 3), ("f32", 16, 16),
                                    ("f32", 12, 4), ("f32", 1, 64), ("bf16", 8, 8),
                                    ("bf16", 64, 16)])
def test_tsmm_against_oracle_and_backward(dt, k, m):
    n = 20011
    A = rand(n, k, DT[dt], 1)
    xdt = torch.float64 if dt == "f64" else torch.float32
    X = torch.from_numpy(oracle.gen_matrix(k, m, dist=oracle.DIST_NORMAL, seed=SEED, stream=3))
    Y = ops.tsmm(A, X.to("cuda", xdt))
    # closed form on the host: dB = A dC with dC = X
    _, ref = oracle.atb_backward(host(A), np.zeros((n, m)), X.to(xdt).double().numpy())
    if dt == "f64":
        tol = dict(rtol=1e-13, atol=1e-13)
    elif dt == "f32":
        tol = dict(rtol=1e-5, atol=1e-5)
    else:
        tol = dict(rtol=2**-7, atol=2**-7 * np.abs(X.numpy()).sum())
    np.testing.assert_allclose(host(Y), ref, **tol)
    # the same kernels as the backward pass
    B = torch.zeros((n, m), dtype=DT[dt], device="cuda")
    _, dB = ops.atb_backward(A, B, X.to("cuda", xdt), need_dA=False)
    assert torch.equal(Y, dB)
    # a strided view of A takes the row-wise kernel: the same bits as the
    # staged kernel (widths up to 8), within the tolerance for the
    # matrix-core classes
    W = torch.zeros((n, k + 2), dtype=DT[dt], device="cuda")
    W[:, :k] = A
    Yw = ops.tsmm(W[:, :k], X.to("cuda", xdt))
    if max(k, m) <= 8:
        assert torch.equal(Yw, Y)
    else:
        np.testing.assert_allclose(host(Yw), ref, **tol)


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_tsmm_bf16_rounding_is_nearest_even(w):
    """With a diagonal X every output is one FP32 product a * x rounded once
    to BF16 (adding the FP32 zero products is exact), so the outputs of the
    packed (w <= 4) and staged kernels must equal torch's CPU conversion
    (round to nearest even) of the FP32 product, bit for bit; a strided view
    (row-wise kernel) must give the same bits.  (Wider classes run on matrix
    instructions, whose FP32 sums are not rounded to nearest.)"""
    n = 50021
    A = rand(n, w, torch.bfloat16, 11)
    rng = np.random.default_rng(SEED + w)
    d = (rng.uniform(0.5, 2.0, w) * np.where(rng.random(w) < 0.5, -1.0, 1.0)).astype(np.float32)
    X = torch.diag(torch.from_numpy(d))
    Y = ops.tsmm(A, X.cuda())
    prod = A.cpu().float() * torch.from_numpy(d)[None, :]  # FP32 products, rounded once
    ref = prod.to(torch.bfloat16)
    assert torch.equal(Y.cpu().view(torch.int16), ref.view(torch.int16))
    W = torch.zeros((n, w + 3), dtype=torch.bfloat16, device="cuda")
    W[:, :w] = A
    Yw = ops.tsmm(W[:, :w], X.cuda())
    assert torch.equal(Yw.cpu().view(torch.int16), ref.view(torch.int16))

