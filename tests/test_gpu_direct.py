"""GPU parity of the direct matrix-core engine (engine 3; run on an MI355X: -m gpu).

The engine streams each wave's rows straight into matrix-instruction
operands (FP64, FP32) or through a wave-private LDS tile (BF16).  Every
direct kernel of the library is checked against the oracle with the
tolerances of test_gpu_parity.py:
  FP64:       |C - C_exact| <= 4 sqrt(n) 2^-52 (|A|^T |B|)_ij
  FP32/BF16:  the same form with 2^-23 and Frobenius relative error <= 1e-6.
Row counts cover the pipelined steps, the steps dealt one at a time after
them, and the n % 4 final rows; requests for the engine on inputs it cannot
read (strided rows, k or m below the class width) run the class's staged
kernel and must give the same bound.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import _lib, ops  # noqa: E402

SEED = oracle.SEED
DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}
EPS = {"f64": 2.0**-52, "f32": 2.0**-23, "bf16": 2.0**-23}
DID = {"f64": _lib.F64, "f32": _lib.F32, "bf16": _lib.BF16}
CLASSES = [("f64", 16, 16), ("f64", 32, 32), ("f64", 64, 64), ("f64", 16, 64), ("f64", 64, 32),
           ("f32", 16, 16), ("f32", 32, 32), ("f32", 64, 64), ("f32", 32, 64),
           ("bf16", 16, 16), ("bf16", 32, 32), ("bf16", 64, 64), ("bf16", 16, 64)]
# 100003 = 25000 steps + 3 rows: with 256 workgroups and 8 steps in flight,
# 24 pipelined steps per wave, 424 dealt one at a time, 3 final rows (BF16:
# 32-row steps, so 3125 steps + 3 rows)
NS = [0, 1, 3, 4, 5, 31, 33, 127, 128, 1001, 100003]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    torch.cuda.set_device(0)
    _lib.lib()
    yield


def rand(n, w, dt, stream, dist=0):
    return ops.random_matrix(n, w, DT[dt], seed=SEED, stream_id=stream, dist=dist)


def host(X):
    return X.cpu().double().numpy()


def check(C, Ah, Bh, dt):
    n = Ah.shape[0]
    ref, absC = oracle.atb_dot2(Ah, Bh, with_abs=True)
    bound = 4.0 * np.sqrt(max(n, 1)) * EPS[dt] * absC
    err = np.abs(C - ref)
    assert np.all(err <= bound), f"max err/bound {np.max(err / np.maximum(bound, 1e-300)):.3g}"
    if dt != "f64" and np.linalg.norm(ref) > 0:
        assert np.linalg.norm(C - ref) / np.linalg.norm(ref) <= 1e-6


def direct_kernels(dt, kc, mc):
    ks = [kd for kd in _lib.kernels() if kd["engine"] == _lib.ENGINE_DIRECT and
          kd["in_type"] == DID[dt] and kd["kc"] == kc and kd["mc"] == mc]
    assert ks, f"no direct kernel for {dt} {kc}x{mc}"
    return ks


@pytest.mark.parametrize("cls", CLASSES, ids=lambda c: f"{c[0]}-{c[1]}x{c[2]}")
def test_direct_against_oracle(cls):
    dt, kc, mc = cls
    for kd in direct_kernels(dt, kc, mc):
        cfg = _lib.config_of(kd)
        for n in NS:
            print(f"{kd['name']} n={n}", flush=True)
            A, B = rand(n, kc, dt, 1), rand(n, mc, dt, 2)
            idx, _ = _lib.plan(DID[dt], n, kc, mc, cfg)
            assert idx == kd["index"]
            C = ops.atb(A, B, out_dtype=torch.float64, config=cfg).cpu().numpy()
            if n == 0:
                assert np.array_equal(C, np.zeros((kc, mc)))
                continue
            check(C, host(A), host(B), dt)


@pytest.mark.parametrize("cls", [("f64", 64, 64), ("f32", 32, 32), ("f64", 16, 16), ("bf16", 64, 64)],
                         ids=lambda c: f"{c[0]}-{c[1]}x{c[2]}")
@pytest.mark.parametrize("grid", [1, 3, 77])
def test_direct_grids_and_pow2(cls, grid):
    """Small forced grids (many pipelined steps per wave, uneven leftovers) on
    sign * 2^e inputs (spread over 41 binades)."""
    dt, kc, mc = cls
    n = 54321
    A, B = rand(n, kc, dt, 3, 1), rand(n, mc, dt, 4, 1)
    for kd in direct_kernels(dt, kc, mc):
        print(f"{kd['name']} grid={grid}", flush=True)
        C = ops.atb(A, B, out_dtype=torch.float64, config=_lib.config_of(kd, grid=grid))
        check(C.cpu().numpy(), host(A), host(B), dt)


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_direct_request_on_inputs_it_cannot_read(dt):
    """Strided rows and widths below the class run the staged fallback."""
    cfg = _lib.make_config(engine=_lib.ENGINE_DIRECT)
    n = 20011
    W = rand(n, 67, dt, 5)
    for k, m in ((64, 64), (57, 64), (64, 40), (33, 48)):
        A, B = W[:, :k], W[:, 3:3 + m]
        C = ops.atb(A, B, out_dtype=torch.float64, config=cfg).cpu().numpy()
        check(C, host(A), host(B), dt)
        Ad, Bd = A.contiguous(), B.contiguous()
        C2 = ops.atb(Ad, Bd, out_dtype=torch.float64, config=cfg).cpu().numpy()
        check(C2, host(Ad), host(Bd), dt)


def test_direct_deterministic_and_accumulate():
    n = 1 << 21
    A, B = rand(n, 64, "f64", 6), rand(n, 64, "f64", 7)
    cfg = _lib.make_config(engine=_lib.ENGINE_DIRECT)
    first = ops.atb(A, B, config=cfg)
    for _ in range(3):
        assert torch.equal(ops.atb(A, B, config=cfg), first)
    C0 = torch.from_numpy(oracle.gen_matrix(64, 64, seed=SEED, stream=8)).cuda()
    C = C0.clone()
    ops.atb(A, B, out=C, accumulate=True, config=cfg)
    assert torch.equal(C, C0 + first)
    C32 = ops.atb(A, B, out_dtype=torch.float32, config=cfg)
    assert torch.equal(C32, first.float())


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_direct_ata(dt):
    """C = A^T A on the direct engine (blocks on and above the diagonal,
    mirrored): the oracle bound, exact symmetry, and the staged fallback for
    strided rows."""
    cfg = _lib.make_config(engine=_lib.ENGINE_DIRECT)
    for n in (1, 6, 1001, 100003):
        print(f"ata {dt} n={n}", flush=True)
        A = rand(n, 64, dt, 9)
        C = ops.ata(A, out_dtype=torch.float64, config=cfg).cpu().numpy()
        Ah = host(A)
        check(C, Ah, Ah, dt)
        assert np.array_equal(C, C.T)
    W = rand(30011, 66, dt, 10)
    C = ops.ata(W[:, 1:65], out_dtype=torch.float64, config=cfg).cpu().numpy()
    Ah = host(W[:, 1:65])
    check(C, Ah, Ah, dt)
    assert np.array_equal(C, C.T)


def test_direct_64bit_offsets():
    """Byte offsets past 2^34 (FP32 64 x 64, n = 2^26 + 5): the whole equals
    the sum of its two halves (split-n additivity) to the FP32 partials'
    rounding."""
    n = (1 << 26) + 5
    A, B = rand(n, 64, "f32", 11), rand(n, 64, "f32", 12)
    cfg = _lib.make_config(engine=_lib.ENGINE_DIRECT)
    C = ops.atb(A, B, out_dtype=torch.float64, config=cfg)
    h = n // 2 + 3
    C1 = ops.atb(A[:h], B[:h], out_dtype=torch.float64, config=cfg)
    C2 = ops.atb(A[h:], B[h:], out_dtype=torch.float64, config=cfg)
    ref = C1 + C2
    # both sides carry FP32-accumulated 64-row partials: relative 1e-6 of the scale
    assert torch.allclose(C, ref, rtol=1e-6, atol=1e-6 * ref.abs().max().item())
    del A, B
