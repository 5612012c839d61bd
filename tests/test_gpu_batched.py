"""GPU parity of the batched entry points (tsm_atb_batched, the plan API)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import _lib, ops  # noqa: E402

SEED = oracle.SEED
EPS = {torch.float64: 2.0**-52, torch.float32: 2.0**-23, torch.bfloat16: 2.0**-23}


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    torch.cuda.set_device(0)
    yield


def host(X):
    if X.dtype == torch.bfloat16:
        return oracle.bf16_bits_to_f64(X.cpu().view(torch.int16).numpy().view(np.uint16))
    return X.cpu().double().numpy()


def make_batch(dt, count, widths, nmax_log2, seed):
    rng = np.random.default_rng(seed)
    As, Bs = [], []
    for i in range(count):
        n = int(round(2 ** rng.uniform(0, nmax_log2))) if i % 7 else int(rng.integers(0, 3))
        k, m = int(rng.choice(widths)), int(rng.choice(widths))
        As.append(ops.random_matrix(n, k, dt, seed=SEED, stream_id=1000 + 2 * i))
        Bs.append(ops.random_matrix(n, m, dt, seed=SEED, stream_id=1001 + 2 * i))
    return As, Bs


def check(C, A, B, dt):
    Ah, Bh = host(A), host(B)
    n = Ah.shape[0]
    if n == 0:
        assert np.array_equal(C, np.zeros_like(C))
        return
    ref, absC = oracle.atb_dot2(Ah, Bh, with_abs=True)
    assert np.all(np.abs(C - ref) <= 4 * np.sqrt(n) * EPS[dt] * absC)


@pytest.mark.parametrize("dt,widths", [
    (torch.float64, [1, 2, 3, 4, 5, 8, 13, 16, 32, 33, 64]),
    (torch.float32, [1, 2, 3, 4, 7, 8, 16, 17, 32]),
    (torch.bfloat16, [2, 4, 8, 16, 32]),
])
def test_batched_against_oracle(dt, widths):
    As, Bs = make_batch(dt, 80, widths, 17, 1)
    outs = ops.atb_batched(As, Bs, out_dtype=torch.float64)
    again = ops.atb_batched(As, Bs, out_dtype=torch.float64)
    for A, B, C, C2 in zip(As, Bs, outs, again):
        assert torch.equal(C, C2), "batched results must be deterministic"
        check(C.cpu().numpy(), A, B, dt)


def test_batched_accumulate_f32_out_and_views():
    As, Bs = make_batch(torch.float64, 30, [2, 4, 8], 14, 2)
    As = [A[:, : max(1, A.shape[1] - 1)] for A in As]  # strided views (lda > k)
    base = [torch.full((A.shape[1], B.shape[1]), 0.5, dtype=torch.float32, device="cuda")
            for A, B in zip(As, Bs)]
    outs = [b.clone() for b in base]
    ops.atb_batched(As, Bs, outs=outs, accumulate=True)
    for A, B, C in zip(As, Bs, outs):
        ref = oracle.atb_dot2(host(A), host(B)) + 0.5
        np.testing.assert_allclose(C.cpu().numpy(), ref, rtol=2e-7, atol=2e-7 * (1 + np.abs(ref).max()))


def test_batched_matches_single_calls_at_the_batched_workload_shape():
    """512 pairs, n log-uniform in [2^10, 2^21], k, m in {2, 4, 8, 16, 32}."""
    rng = np.random.default_rng(20261203)
    As, Bs = [], []
    for i in range(512):
        n = int(round(2 ** rng.uniform(10, 21)))
        k, m = (int(x) for x in rng.choice([2, 4, 8, 16, 32], 2))
        As.append(ops.random_matrix(n, k, torch.float64, seed=SEED, stream_id=3 * i + 1))
        Bs.append(ops.random_matrix(n, m, torch.float64, seed=SEED, stream_id=3 * i + 2))
    outs = ops.atb_batched(As, Bs)
    for A, B, C in zip(As, Bs, outs):
        single = ops.atb(A, B)
        rel = ((C - single).norm() / single.norm()).item()
        assert rel <= 1e-13, rel


def test_batched_large_problem_with_capped_items():
    """A problem that would be cut into more than 256 items of 4 MiB runs as
    256 larger items (next to a small problem): the same results as a single
    call within 1e-13 relative, deterministic."""
    A = ops.random_matrix((1 << 21) + 77, 32, torch.float64, seed=SEED, stream_id=1)  # 1 GiB + A
    B = ops.random_matrix((1 << 21) + 77, 32, torch.float64, seed=SEED, stream_id=2)
    a2 = ops.random_matrix(1001, 8, torch.float64, seed=SEED, stream_id=3)
    b2 = ops.random_matrix(1001, 4, torch.float64, seed=SEED, stream_id=4)
    plan_outs = [torch.empty((32, 32), dtype=torch.float64, device="cuda"),
                 torch.empty((8, 4), dtype=torch.float64, device="cuda")]
    plan = ops.BatchedPlan([A, a2], [B, b2], plan_outs)
    assert plan.items <= 256 + 1
    plan.run()
    torch.cuda.synchronize()
    for C, X, Y in zip(plan_outs, (A, a2), (B, b2)):
        single = ops.atb(X, Y)
        assert ((C - single).norm() / single.norm()).item() <= 1e-13
    first = [C.clone() for C in plan_outs]
    plan.run()
    torch.cuda.synchronize()
    assert all(torch.equal(C, F) for C, F in zip(plan_outs, first))
    plan.close()


def test_queue_deals_items_beyond_the_grid():
    """Items past the first grid-full come from the kernel's queue, whose
    tickets the kernel zeroes at its end: many more items than workgroups, run
    eagerly and by graph replay several times, give the same bits each time
    and match single calls."""
    rng = np.random.default_rng(7)
    As, Bs = [], []
    for i in range(48):
        n = int(rng.integers(1 << 19, 1 << 20))
        k, m = (int(x) for x in rng.choice([2, 8, 16, 32], 2))
        As.append(ops.random_matrix(n, k, torch.float64, seed=SEED, stream_id=5000 + 2 * i))
        Bs.append(ops.random_matrix(n, m, torch.float64, seed=SEED, stream_id=5001 + 2 * i))
    outs = [torch.empty((A.shape[1], B.shape[1]), dtype=torch.float64, device="cuda")
            for A, B in zip(As, Bs)]
    plan = ops.BatchedPlan(As, Bs, outs)
    assert plan.items > 4 * plan.grid, (plan.items, plan.grid)
    plan.run()
    torch.cuda.synchronize()
    first = [C.clone() for C in outs]
    for A, B, C in zip(As, Bs, first):
        single = ops.atb(A, B)
        assert ((C - single).norm() / single.norm()).item() <= 1e-13
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run()
    for i in range(3):
        for C in outs:
            C.zero_()
        plan.run() if i % 2 else g.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(C, F) for C, F in zip(outs, first)), f"call {i}"
    plan.close()


def test_plan_api_and_graph_replay():
    As, Bs = make_batch(torch.float32, 40, [2, 4, 8, 16, 32], 16, 3)
    outs = [torch.empty((A.shape[1], B.shape[1]), dtype=torch.float64, device="cuda")
            for A, B in zip(As, Bs)]
    plan = ops.BatchedPlan(As, Bs, outs)
    assert plan.items >= len(As) and plan.grid >= 1
    plan.run()
    torch.cuda.synchronize()
    eager = [C.clone() for C in outs]
    for A, B, C in zip(As, Bs, eager):
        check(C.cpu().numpy(), A, B, torch.float32)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run()
    for i in range(4):
        for C in outs:
            C.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(C, E) for C, E in zip(outs, eager)), f"replay {i}"
        plan.run()  # eager execution between replays
        torch.cuda.synchronize()
        assert all(torch.equal(C, E) for C, E in zip(outs, eager))
    plan.close()


@pytest.mark.parametrize("dt", [torch.float32, torch.bfloat16])
def test_batched_wide_low_precision_as_column_blocks(dt):
    """FP32/BF16 problems with k or m above 32 run as column blocks of at most
    32 x 32 in the same launch: each C against the oracle, deterministic, and
    accumulate / FP32 output / strided views on the split problems too."""
    rng = np.random.default_rng(SEED + 7)
    As, Bs = [], []
    for i, (k, m) in enumerate([(64, 64), (33, 8), (8, 40), (48, 64), (16, 16), (64, 1), (2, 33)]):
        n = int(rng.integers(1, 40000)) if i else 0
        As.append(ops.random_matrix(n, k, dt, seed=SEED, stream_id=2 * i + 1))
        Bs.append(ops.random_matrix(n, m, dt, seed=SEED, stream_id=2 * i + 2))
    outs = ops.atb_batched(As, Bs, out_dtype=torch.float64)
    again = ops.atb_batched(As, Bs, out_dtype=torch.float64)
    for A, B, C, C2 in zip(As, Bs, outs, again):
        assert torch.equal(C, C2)
        check(C.cpu().numpy(), A, B, dt)
    # the plan API and a graph replay of it give the same bits
    pouts = [torch.empty_like(C) for C in outs]
    plan = ops.BatchedPlan(As, Bs, pouts)
    plan.run()
    torch.cuda.synchronize()
    assert all(torch.equal(C, P) for C, P in zip(outs, pouts))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run()
    for P in pouts:
        P.zero_()
    g.replay()
    torch.cuda.synchronize()
    assert all(torch.equal(C, P) for C, P in zip(outs, pouts))
    plan.close()
    # views with lda > k, C += A^T B into FP32 outputs
    W = ops.random_matrix(30011, 67, dt, seed=SEED, stream_id=40)
    As2, Bs2 = [W[:, :64], W[:, 2:50]], [W[:, 3:67], W[:, :40]]
    base = [torch.full((A.shape[1], B.shape[1]), 0.25, dtype=torch.float32, device="cuda")
            for A, B in zip(As2, Bs2)]
    outs2 = [b.clone() for b in base]
    ops.atb_batched(As2, Bs2, outs=outs2, accumulate=True)
    for A, B, C in zip(As2, Bs2, outs2):
        ref = oracle.atb_dot2(host(A), host(B)) + 0.25
        np.testing.assert_allclose(C.cpu().numpy(), ref, rtol=2e-6, atol=2e-6 * (1 + np.abs(ref).max()))
