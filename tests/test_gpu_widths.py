"""Every width pair and every short row count on the GPU (run on an MI355X:
-m gpu).

The library maps each (k, m) with k, m <= 64 onto a generated class (exact,
padded or split into column blocks) and each backward width onto one of its
kernels, so every pair takes a mapping path of its own.  These tests run all
4096 pairs of C = A^T B per storage type, dense and as strided views, in
single calls and in one batched launch, every A^T A width, and both
gradients of every pair, each against the oracle.  Then every n up to 520
and every n within two rows of one to three tiles, for forward classes of
each engine and backward classes of each kernel.

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
from tsm_benchmark_amd import _lib, ops  # noqa: E402

SEED = oracle.SEED
W = 64
DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}
EPS = {"f64": 2.0**-52, "f32": 2.0**-23, "bf16": 2.0**-23}
IN = {"f64": _lib.F64, "f32": _lib.F32, "bf16": _lib.BF16}
GUARD = 7.0  # fill of the guard rows around written outputs (exact in every type)
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
    L = _lib.lib()
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
        # the dense kernels again, through the C ABI, writing dA and dB into
        # the middle of buffers with 16 guard rows on either side
        gA = torch.full((n + 32, k), GUARD, dtype=DT[dt], device="cuda")
        gB = torch.full((n + 32, m), GUARD, dtype=DT[dt], device="cuda")
        G_km = Gd[:k, :m].contiguous()
        st = L.tsm_atb_backward(IN[dt], IN[dt if dt == "f64" else "f32"], n, k, m,
                                A_k[k].data_ptr(), k, B_m[m].data_ptr(), m, G_km.data_ptr(), m,
                                gA[16:].data_ptr(), k, gB[16:].data_ptr(), m,
                                torch.cuda.current_stream().cuda_stream)
        assert st == _lib.SUCCESS, (k, m, st)
        dA, dB = gA[16:16 + n], gB[16:16 + n]
        ea = ((dA.double() - ra).abs() - rtol * ra.abs() - atol).max()
        eb = ((dB.double() - rb).abs() - rtol * rb.abs() - atol).max()
        spill = ((gA[:16] != GUARD).any() | (gA[16 + n:] != GUARD).any() |
                 (gB[:16] != GUARD).any() | (gB[16 + n:] != GUARD).any())
        e = torch.where(spill, torch.full_like(ea, float("inf")), torch.maximum(ea, eb))
        excess.append(torch.maximum(worst, e))
    excess = torch.stack(excess).cpu().numpy()
    bad = [PAIRS[i] for i in np.flatnonzero(excess > 0)]
    assert not bad, f"{len(bad)} failing (k, m), first: {bad[:8]}, worst excess {excess.max():.3g}"


# ---- every short row count -----------------------------------------------------------------
# (storage, k, m, tile rows of the tuned forward kernel): VALU register tiles,
# LDS-staged and direct matrix-core kernels, tiles of 4 to 4096 rows
FWD_CLASSES = [("f64", 8, 8, 128), ("f64", 1, 1, 4096), ("f64", 32, 32, 32), ("f64", 64, 64, 4),
               ("f32", 4, 4, 1024), ("f32", 16, 16, 128), ("f32", 64, 64, 4),
               ("bf16", 1, 1, 4096), ("bf16", 32, 32, 512), ("bf16", 64, 64, 32)]


def row_counts(tile):
    """Every n up to 520, and n within 2 of one, two and three tiles."""
    ns = set(range(521))
    for j in (1, 2, 3):
        ns.update(j * tile + d for d in range(-2, 3))
    return sorted(ns)


@pytest.mark.parametrize("cls", FWD_CLASSES, ids=lambda c: f"{c[0]}-{c[1]}x{c[2]}")
def test_every_short_row_count_forward(cls):
    """Each tail length of the dense kernels: n rows are a prefix of one
    matrix pair, one launch per n, each against the oracle on that prefix."""
    dt, k, m, tile = cls
    ns = row_counts(tile)
    A = ops.random_matrix(ns[-1], k, DT[dt], seed=SEED, stream_id=1)
    B = ops.random_matrix(ns[-1], m, DT[dt], seed=SEED, stream_id=2)
    out = [ops.atb(A[:n], B[:n], out_dtype=torch.float64) for n in ns]
    torch.cuda.synchronize()
    Ah, Bh = host(A), host(B)
    bad = []
    for n, C in zip(ns, out):
        C = C.cpu().numpy()
        if n == 0:
            ok = np.array_equal(C, np.zeros((k, m)))
        else:
            r, absC = oracle.atb_dot2(Ah[:n], Bh[:n], with_abs=True)
            ok = np.all(np.abs(C - r) <= 4.0 * np.sqrt(n) * EPS[dt] * absC)
            if ok and dt != "f64" and np.linalg.norm(r) > 0:
                ok = np.linalg.norm(C - r) / np.linalg.norm(r) <= 1e-6
        if not ok:
            bad.append(n)
    assert not bad, f"{len(bad)} failing row counts, first: {bad[:10]}"


# packed (rows of at most 16 bytes), LDS-staged and matrix-core backward
# kernels; their tiles are 128 or 256 rows
BWD_CLASSES = [("f64", 2, 2), ("f64", 4, 4), ("f64", 8, 8), ("f64", 16, 16), ("f64", 64, 64),
               ("f32", 4, 4), ("f32", 16, 16), ("f32", 32, 32), ("bf16", 8, 8), ("bf16", 32, 32),
               ("bf16", 64, 16)]


@pytest.mark.parametrize("cls", BWD_CLASSES, ids=lambda c: f"{c[0]}-{c[1]}x{c[2]}")
def test_every_short_row_count_backward(cls):
    """Each tail length of the backward kernels.  Row r of dA (dB) depends on
    row r of B (A) only, so the oracle at the largest n serves every prefix."""
    dt, k, m = cls
    ns = row_counts(256)
    N = ns[-1]
    gdt = torch.float64 if dt == "f64" else torch.float32
    A = ops.random_matrix(N, k, DT[dt], seed=SEED, stream_id=1)
    B = ops.random_matrix(N, m, DT[dt], seed=SEED, stream_id=2)
    G = torch.from_numpy(oracle.gen_matrix(k, m, dist=oracle.DIST_NORMAL, seed=SEED, stream=3)).to(gdt)
    dA_ref, dB_ref = oracle.atb_backward(host(A), host(B), G.double().numpy())
    dA_ref, dB_ref = torch.from_numpy(dA_ref).to("cuda"), torch.from_numpy(dB_ref).to("cuda")
    if dt == "f64":
        rtol = atol = 1e-13
    elif dt == "f32":
        rtol = atol = 1e-5
    else:
        rtol, atol = 2.0**-7, 2.0**-7 * float(np.abs(G.double().numpy()).sum())
    Gd = G.to("cuda")
    excess = []
    for n in ns:
        dA, dB = ops.atb_backward(A[:n], B[:n], Gd)
        if n == 0:
            continue
        ra, rb = dA_ref[:n], dB_ref[:n]
        ea = ((dA.double() - ra).abs() - rtol * ra.abs() - atol).max()
        eb = ((dB.double() - rb).abs() - rtol * rb.abs() - atol).max()
        excess.append(torch.maximum(ea, eb))
    excess = torch.stack(excess).cpu().numpy()
    bad = [n for n, e in zip(ns[1:], excess) if e > 0]
    assert not bad, f"{len(bad)} failing row counts, first: {bad[:10]}"


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("op", ["atb", "ata"])
def test_every_width_accumulates_into_a_strided_fp32_output(dt, op):
    """C += A^T B (or A^T A) in FP32 through a k x m view of a 64 x 64 buffer
    (ldc = 64) for every width: each view ends as fl32(C0 + sum), and no
    element of the buffer outside the view changes."""
    n = 4099
    A, B = rand(n, dt, 1), rand(n, dt, 2)
    Ah = host(A)
    Bh = Ah if op == "ata" else host(B)
    ref, absC = oracle.atb_dot2(Ah, Bh, with_abs=True)
    bound = 4.0 * np.sqrt(n) * EPS[dt] * absC
    A_k = [None] + [A[:, :k].contiguous() for k in range(1, W + 1)]
    B_m = [None] + [B[:, :m].contiguous() for m in range(1, W + 1)]
    pairs = [(k, k) for k in range(1, W + 1)] if op == "ata" else PAIRS
    g = torch.Generator(device="cuda").manual_seed(7)
    bufs = []
    for k, m in pairs:
        buf = torch.rand((W, W), dtype=torch.float32, device="cuda", generator=g) * 4 - 2
        before = buf.clone()
        C = buf[:k, :m]
        if op == "ata":
            ops.ata(A_k[k], out=C, accumulate=True)
        else:
            ops.atb(A_k[k], B_m[m], out=C, accumulate=True)
        bufs.append((buf, before))
    torch.cuda.synchronize()
    bad = []
    for (k, m), (buf, before) in zip(pairs, bufs):
        got, b0 = buf.cpu().double().numpy(), before.cpu().double().numpy()
        want = b0[:k, :m] + ref[:k, :m]
        tol = bound[:k, :m] + 2.0**-23 * np.abs(want) + 2.0**-24 * np.abs(b0[:k, :m])
        inside = np.all(np.abs(got[:k, :m] - want) <= tol)
        rest = got.copy()
        rest[:k, :m] = b0[:k, :m]
        if not (inside and np.array_equal(rest, b0)):
            bad.append((k, m, bool(inside)))
    assert not bad, f"{len(bad)} failing (k, m, inside ok), first: {bad[:8]}"


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
def test_every_width_pair_in_one_batch_accumulates_into_strided_fp32_outputs(dt):
    """The 4096 pairs of one batched launch, each adding into an FP32 k x m
    view of its own 64 x 64 buffer (ldc = 64): each view ends as
    fl32(C0 + sum), and no element outside the views changes."""
    n = 4099
    A, B = rand(n, dt, 1), rand(n, dt, 2)
    ref, absC = oracle.atb_dot2(host(A), host(B), with_abs=True)
    bound = 4.0 * np.sqrt(n) * EPS[dt] * absC
    A_k = [None] + [A[:, :k].contiguous() for k in range(1, W + 1)]
    B_m = [None] + [B[:, :m].contiguous() for m in range(1, W + 1)]
    g = torch.Generator(device="cuda").manual_seed(11)
    bufs = torch.rand((len(PAIRS), W, W), dtype=torch.float32, device="cuda", generator=g) * 4 - 2
    before = bufs.clone()
    outs = [bufs[p, :k, :m] for p, (k, m) in enumerate(PAIRS)]
    ops.atb_batched([A_k[k] for k, _ in PAIRS], [B_m[m] for _, m in PAIRS], outs=outs,
                    accumulate=True)
    got, b0 = bufs.cpu().double().numpy(), before.cpu().double().numpy()
    bad = []
    for p, (k, m) in enumerate(PAIRS):
        want = b0[p, :k, :m] + ref[:k, :m]
        tol = bound[:k, :m] + 2.0**-23 * np.abs(want) + 2.0**-24 * np.abs(b0[p, :k, :m])
        inside = np.all(np.abs(got[p, :k, :m] - want) <= tol)
        rest = got[p].copy()
        rest[:k, :m] = b0[p, :k, :m]
        if not (inside and np.array_equal(rest, b0[p])):
            bad.append((k, m, bool(inside)))
    assert not bad, f"{len(bad)} failing (k, m, inside ok), first: {bad[:8]}"


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16"])
def test_every_width_on_misaligned_views(dt):
    """Views that start one element into a 69-column buffer: every base
    misaligned by one element, ld = 69.  A^T B for all 4096 pairs, A^T A for
    every width and both gradients for every pair, against the oracle."""
    n, ld, off = 4099, W + 5, 1
    A = ops.random_matrix(n, ld, DT[dt], seed=SEED, stream_id=5)
    B = ops.random_matrix(n, ld, DT[dt], seed=SEED, stream_id=6)
    Ah, Bh = host(A)[:, off:off + W], host(B)[:, off:off + W]
    ref, absC = oracle.atb_dot2(Ah, Bh, with_abs=True)
    bound = 4.0 * np.sqrt(n) * EPS[dt] * absC
    sref, sabs = oracle.atb_dot2(Ah, Ah, with_abs=True)
    sbound = 4.0 * np.sqrt(n) * EPS[dt] * sabs
    Av = [None] + [A[:, off:off + k] for k in range(1, W + 1)]
    Bv = [None] + [B[:, off:off + m] for m in range(1, W + 1)]
    fwd = [ops.atb(Av[k], Bv[m], out_dtype=torch.float64) for k, m in PAIRS]
    sym = [ops.ata(Av[k], out_dtype=torch.float64) for k in range(1, W + 1)]
    # gradients: one oracle call per width at the full other width, sliced
    gdt = torch.float64 if dt == "f64" else torch.float32
    G = torch.from_numpy(oracle.gen_matrix(W, W, dist=oracle.DIST_NORMAL, seed=SEED, stream=3)).to(gdt)
    Gh, Gd = G.double().numpy(), G.to("cuda")
    dA_ref = [None] + [torch.from_numpy(oracle.atb_backward(Ah, Bh[:, :m], Gh[:, :m])[0]).to("cuda")
                       for m in range(1, W + 1)]
    dB_ref = [None] + [torch.from_numpy(oracle.atb_backward(Ah[:, :k], Bh, Gh[:k, :])[1]).to("cuda")
                       for k in range(1, W + 1)]
    absG = np.abs(Gh)
    excess = []
    for k, m in PAIRS:
        if dt == "f64":
            rtol = atol = 1e-13
        elif dt == "f32":
            rtol = atol = 1e-5
        else:
            rtol, atol = 2.0**-7, 2.0**-7 * float(absG[:k, :m].sum())
        dA, dB = ops.atb_backward(Av[k], Bv[m], Gd[:k, :m])
        ra, rb = dA_ref[m][:, :k], dB_ref[k][:, :m]
        excess.append(torch.maximum(((dA.double() - ra).abs() - rtol * ra.abs() - atol).max(),
                                    ((dB.double() - rb).abs() - rtol * rb.abs() - atol).max()))
    excess = torch.stack(excess).cpu().numpy()
    bad = [("bwd",) + PAIRS[i] for i in np.flatnonzero(excess > 0)]
    for (k, m), C in zip(PAIRS, fwd):
        C = C.cpu().numpy()
        ok = np.all(np.abs(C - ref[:k, :m]) <= bound[:k, :m])
        if ok and dt != "f64" and np.linalg.norm(ref[:k, :m]) > 0:
            ok = np.linalg.norm(C - ref[:k, :m]) / np.linalg.norm(ref[:k, :m]) <= 1e-6
        if not ok:
            bad.append(("atb", k, m))
    for k, C in zip(range(1, W + 1), sym):
        C = C.cpu().numpy()
        if not (np.all(np.abs(C - sref[:k, :k]) <= sbound[:k, :k]) and np.array_equal(C, C.T)):
            bad.append(("ata", k, k))
    assert not bad, f"{len(bad)} failing (op, k, m), first: {bad[:8]}"
