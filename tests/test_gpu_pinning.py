"""Checks that pin the GPU path where the other parity tests leave gaps: the
FP64 central-difference check of the backward pass at the differentiation
workload's check size (n = 4096, k = m = 4, h = 1e-6), and the FP64 forward
on sign * 2^e inputs against the exact int128 oracle across the width classes."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import ops  # noqa: E402

SEED = oracle.SEED


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    torch.cuda.set_device(0)
    yield


def test_backward_central_differences_at_n4096_k4_m4():
    """L(A, B) = <G, A^T B> is bilinear, so a central difference with h = 1e-6
    equals the derivative up to rounding; both L and the gradients come from
    the GPU (forward kernel, hand-written backward kernels)."""
    n, k, m, h = 4096, 4, 4, 1e-6
    A0 = torch.from_numpy(oracle.gen_matrix(n, k, seed=SEED, stream=21)).cuda()
    B0 = torch.from_numpy(oracle.gen_matrix(n, m, seed=SEED, stream=22)).cuda()
    G = torch.from_numpy(oracle.gen_matrix(k, m, dist=oracle.DIST_NORMAL, seed=SEED, stream=23)).cuda()
    A, B = A0.clone().requires_grad_(), B0.clone().requires_grad_()
    (ops.atb_autograd(A, B) * G).sum().backward()
    dA, dB = A.grad.cpu().numpy(), B.grad.cpu().numpy()

    def L(A_, B_):
        return float((ops.atb(A_, B_) * G).sum().item())

    rng = np.random.default_rng(7)
    for r, i in zip(rng.integers(0, n, 12), rng.integers(0, k, 12)):
        for which in (0, 1):
            X = (A0 if which == 0 else B0)
            Xp, Xm = X.clone(), X.clone()
            Xp[r, i] += h
            Xm[r, i] -= h
            step = (Xp[r, i] - Xm[r, i]).item()
            fd = ((L(Xp, B0) - L(Xm, B0)) if which == 0 else (L(A0, Xp) - L(A0, Xm))) / step
            g = (dA if which == 0 else dB)[r, i]
            assert abs(fd - g) <= 1e-6 * (1 + abs(g)), (which, r, i, fd, g)
    # one direction through every entry at once
    VA = torch.from_numpy(oracle.gen_matrix(n, k, seed=SEED, stream=24)).cuda()
    VB = torch.from_numpy(oracle.gen_matrix(n, m, seed=SEED, stream=25)).cuda()
    fd = (L(A0 + h * VA, B0 + h * VB) - L(A0 - h * VA, B0 - h * VB)) / (2 * h)
    want = float((A.grad * VA).sum() + (B.grad * VB).sum())
    assert abs(fd - want) <= 1e-6 * (1 + abs(want)), (fd, want)
    # and the closed form on the host
    np.testing.assert_allclose(dA, B0.cpu().numpy() @ G.cpu().numpy().T, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [1, 2, 8, 32, 64])
def test_pow2_forward_against_the_exact_oracle(k):
    """sign * 2^e entries, e uniform in [-20, 20] (the ill-scaled sweep
    variant): every product and partial sum is a multiple of 2^-40, so the
    int128 oracle gives the exact C; FP64 on the GPU stays within
    4 sqrt(n) 2^-52 (|A|^T |B|) elementwise."""
    n = ((1 << 18) if k == 64 else (1 << 20)) + 3
    A = ops.random_matrix(n, k, torch.float64, seed=SEED, stream_id=31, dist=1)
    B = ops.random_matrix(n, k, torch.float64, seed=SEED, stream_id=32, dist=1)
    C = ops.atb(A, B).cpu().numpy()
    Ah = oracle.gen_matrix(n, k, dist=oracle.DIST_POW2, seed=SEED, stream=31)
    Bh = oracle.gen_matrix(n, k, dist=oracle.DIST_POW2, seed=SEED, stream=32)
    assert np.array_equal(A.cpu().numpy(), Ah) and np.array_equal(B.cpu().numpy(), Bh)
    exact = oracle.atb_exact(Ah, Bh, 20, 20)
    absC = oracle.atb_exact(np.abs(Ah), np.abs(Bh), 20, 20)
    bound = 4.0 * np.sqrt(n) * 2.0**-52 * absC
    err = np.abs(C - exact)
    assert np.all(err <= bound), float(np.max(err / bound))
