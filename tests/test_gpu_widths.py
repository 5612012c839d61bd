"""Every width pair on the GPU (run on an MI355X: -m gpu).

The library maps each (k, m) with k, m <= 64 onto a generated class (exact,
padded or split into column blocks) and each backward width onto one of its
kernels, so every pair takes a mapping path of its own.  These tests run all
4096 pairs of C = A^T B per storage type, dense and as strided views, in
single calls and in one batched launch, every A^T A width, and both
gradients of every pair, each against the oracle.

One oracle call serves every pair: C_ij, dA_ij and dB_ij depend only on the
columns they name, so the 64 x 64 result (or, for the gradients, the result
at the full other width) is sliced.  Tolerances as in the other parity tests:
  forward   |C - C_ref| <= 4 sqrt(n) u (|A|^T |B|)_ij, u = 2^-52 (FP64 inputs)
            or 2^-23 (FP32/BF16 inputs), and a Frobenius relative error
            <= 1e-6 for FP32/BF16 inputs;
  backward  |d - d_ref| <= atol + rtol |d_ref| with rtol = atol = 1e-13
            (FP64), 1e-5 (FP32), rtol = 2^-7 and atol = 2^-7 sum |dC| (BF16).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import ops  # noqa: E402

SEED = oracle.SEED
W = 64
DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}
EPS = {"f64": 2.0**-52, "f32": 2.0**-23, "bf16": 2.0**-23}
# ragged row counts: a partial tile only, a few tiles, a few hundred tiles
NS = (77, 4099, 70001)
PAIRS = [(k, m) for k in range(1, W + 1) for m in range(1, W + 1)]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    torch.cuda.set_device(0)
    yield


def host(X):
    """Device matrix -> FP64 numpy (exact for every supported type)."""
    if X.dtype == torch.bfloat16:
        return oracle.bf16_bits_to_f64(X.cpu().view(torch.int16).numpy().view(np.uint16))
    return X.cpu().double().numpy()


def rand(n, dt, stream):
    return ops.random_matrix(n, W, DT[dt], seed=SEED, stream_id=stream)


def n_of(k, m):
    """Spreads the pairs over the row counts (every n sees every width)."""
    return NS[(7 * k + m) % len(NS)]


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
def test_every_width_pair_forward(dt):
    bad = []
    for n in NS:
        A, B = rand(n, dt, 1), rand(n, dt, 2)
        Ah, Bh = host(A), host(B)
        ref, absC = oracle.atb_dot2(Ah, Bh, with_abs=True)
        bound = 4.0 * np.sqrt(n) * EPS[dt] * absC
        A_k = [None] + [A[:, :k].contiguous() for k in range(1, W + 1)]
        B_m = [None] + [B[:, :m].contiguous() for m in range(1, W + 1)]
        pairs = [(k, m) for k, m in PAIRS if n_of(k, m) == n]
        out = []
        for k, m in pairs:
            dense = ops.atb(A_k[k], B_m[m], out_dtype=torch.float64)
            strided = ops.atb(A[:, :k], B[:, :m], out_dtype=torch.float64)
            out.append((dense, strided))
        torch.cuda.synchronize()
        for (k, m), (dense, strided) in zip(pairs, out):
            r, b = ref[:k, :m], bound[:k, :m]
            for name, C in (("dense", dense), ("strided", strided)):
                C = C.cpu().numpy()
                ok = np.all(np.abs(C - r) <= b)
                if ok and dt != "f64" and np.linalg.norm(r) > 0:
                    ok = np.linalg.norm(C - r) / np.linalg.norm(r) <= 1e-6
                if not ok:
                    bad.append((name, n, k, m))
    assert not bad, f"{len(bad)} failing (path, n, k, m), first: {bad[:8]}"


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
def test_every_width_pair_in_one_batch(dt):
    """All 4096 pairs as the problems of ONE batched launch (the kernel's
    classes, padded widths and column blocks side by side), dense copies and
    strided views."""
    sets = {}
    for n in NS:
        A, B = rand(n, dt, 1), rand(n, dt, 2)
        ref, absC = oracle.atb_dot2(host(A), host(B), with_abs=True)
        sets[n] = (A, B, ref, 4.0 * np.sqrt(n) * EPS[dt] * absC)
    bad = []
    for name in ("dense", "strided"):
        # one copy per (n, width), shared by the pairs that use it
        cols = {}
        for n, (A, B, _, _) in sets.items():
            for w in range(1, W + 1):
                a, b = A[:, :w], B[:, :w]
                cols[n, w] = (a.contiguous(), b.contiguous()) if name == "dense" else (a, b)
        As = [cols[n_of(k, m), k][0] for k, m in PAIRS]
        Bs = [cols[n_of(k, m), m][1] for k, m in PAIRS]
        outs = ops.atb_batched(As, Bs, out_dtype=torch.float64)
        torch.cuda.synchronize()
        for (k, m), C in zip(PAIRS, outs):
            n = n_of(k, m)
            r, b = sets[n][2][:k, :m], sets[n][3][:k, :m]
            C = C.cpu().numpy()
            ok = np.all(np.abs(C - r) <= b)
            if ok and dt != "f64" and np.linalg.norm(r) > 0:
                ok = np.linalg.norm(C - r) / np.linalg.norm(r) <= 1e-6
            if not ok:
                bad.append((name, n, k, m))
        del cols, As, Bs, outs
    assert not bad, f"{len(bad)} failing (path, n, k, m), first: {bad[:8]}"


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
def test_every_width_ata(dt):
    bad = []
    for n in NS:
        A = rand(n, dt, 1)
        Ah = host(A)
        ref, absC = oracle.atb_dot2(Ah, Ah, with_abs=True)
        bound = 4.0 * np.sqrt(n) * EPS[dt] * absC
        out = [(k, ops.ata(A[:, :k].contiguous(), out_dtype=torch.float64),
                ops.ata(A[:, :k], out_dtype=torch.float64)) for k in range(1, W + 1)]
        torch.cuda.synchronize()
        for k, dense, strided in out:
            r, b = ref[:k, :k], bound[:k, :k]
            for name, C in (("dense", dense), ("strided", strided)):
                C = C.cpu().numpy()
                ok = np.all(np.abs(C - r) <= b) and np.array_equal(C, C.T)
                if ok and dt != "f64" and np.linalg.norm(r) > 0:
                    ok = np.linalg.norm(C - r) / np.linalg.norm(r) <= 1e-6
                if not ok:
                    bad.append((name, n, k))
    assert not bad, f"{len(bad)} failing (path, n, k), first: {bad[:8]}"


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
def test_every_width_pair_backward(dt):
    n = 4099
    gdt = torch.float64 if dt == "f64" else torch.float32
    A, B = rand(n, dt, 1), rand(n, dt, 2)
    Ah, Bh = host(A), host(B)
    G = torch.from_numpy(oracle.gen_matrix(W, W, dist=oracle.DIST_NORMAL, seed=SEED, stream=3)).to(gdt)
    Gh = G.double().numpy()
    Gd = G.to("cuda")
    # dA = B dC^T: column i of dA uses row i of dC only, so the k = 64 result
    # at each m holds every k; dB = A dC likewise at m = 64 for each k.
    dA_ref = [None] + [torch.from_numpy(oracle.atb_backward(Ah, Bh[:, :m], Gh[:, :m])[0]).to("cuda")
                       for m in range(1, W + 1)]
    dB_ref = [None] + [torch.from_numpy(oracle.atb_backward(Ah[:, :k], Bh, Gh[:k, :])[1]).to("cuda")
                       for k in range(1, W + 1)]
    A_k = [None] + [A[:, :k].contiguous() for k in range(1, W + 1)]
    B_m = [None] + [B[:, :m].contiguous() for m in range(1, W + 1)]
    absG = np.abs(Gh)
    excess = []  # per pair: max over both gradients and both paths of |d - ref| - rtol |ref| - atol
    for k, m in PAIRS:
        if dt == "f64":
            rtol = atol = 1e-13
        elif dt == "f32":
            rtol = atol = 1e-5
        else:
            rtol, atol = 2.0**-7, 2.0**-7 * float(absG[:k, :m].sum())
        ra, rb = dA_ref[m][:, :k], dB_ref[k][:, :m]
        worst = None
        # dense rows (staged, packed or matrix-core kernels), then strided
        # views (the row-wise kernel)
        for a, b in ((A_k[k], B_m[m]), (A[:, :k], B[:, :m])):
            dA, dB = ops.atb_backward(a, b, Gd[:k, :m])
            ea = ((dA.double() - ra).abs() - rtol * ra.abs() - atol).max()
            eb = ((dB.double() - rb).abs() - rtol * rb.abs() - atol).max()
            e = torch.maximum(ea, eb)
            worst = e if worst is None else torch.maximum(worst, e)
        excess.append(worst)
    excess = torch.stack(excess).cpu().numpy()
    bad = [PAIRS[i] for i in np.flatnonzero(excess > 0)]
    assert not bad, f"{len(bad)} failing (k, m), first: {bad[:8]}, worst excess {excess.max():.3g}"
