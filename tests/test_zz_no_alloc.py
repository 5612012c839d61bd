"""tsm_atb_noalloc (the entry point that allocates nothing) against tsm_atb,
through the C ABI: the same default plan on the same inputs, so the same
bits, on consecutive calls that share the library's static scratch array
(the second call smaller than the first)."""
import ctypes as ct

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import _lib, ops  # noqa: E402

SEED = 20261203
DT = {_lib.F64: torch.float64, _lib.F32: torch.float32, _lib.BF16: torch.bfloat16}


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    torch.cuda.set_device(0)
    yield


def both(dt, n, k, m, sid, accumulate=False):
    """(tsm_atb with a caller workspace, tsm_atb_noalloc) on the same inputs."""
    L = _lib.lib()
    A = ops.random_matrix(n, k, DT[dt], seed=SEED, stream_id=sid)
    B = ops.random_matrix(n, m, DT[dt], seed=SEED, stream_id=sid + 1)
    ws_bytes = L.tsm_atb_workspace_size(dt, n, k, m, None)
    assert 0 < ws_bytes <= _lib.NOALLOC_SCRATCH_BYTES
    ws = torch.zeros(ws_bytes, dtype=torch.uint8, device="cuda")
    C0 = torch.full((k, m), 0.25, dtype=torch.float64, device="cuda")
    C1 = C0.clone()
    s = torch.cuda.current_stream().cuda_stream
    acc = 1 if accumulate else 0
    st = L.tsm_atb(dt, _lib.F64, n, k, m, A.data_ptr(), k, B.data_ptr(), m, C0.data_ptr(), m,
                   acc, ws.data_ptr(), ws_bytes, s)
    assert st == _lib.SUCCESS
    st = L.tsm_atb_noalloc(dt, _lib.F64, n, k, m, A.data_ptr(), k, B.data_ptr(), m,
                           C1.data_ptr(), m, acc, s)
    assert st == _lib.SUCCESS, _lib.strerror(st)
    torch.cuda.synchronize()
    return C0.cpu(), C1.cpu()


@pytest.mark.parametrize("first,second", [
    ((_lib.F64, (1 << 21) + 5, 64, 64), (_lib.F64, 100003, 8, 8)),     # finisher kernel, then in-kernel sum
    ((_lib.F64, (1 << 22), 8, 8), (_lib.F64, 4099, 8, 8)),             # the headline kernel twice
    ((_lib.F32, (1 << 20) + 17, 32, 32), (_lib.BF16, 70001, 16, 16)),  # matrix-core kernels
    ((_lib.F64, 300007, 5, 3), (_lib.F64, 1000, 1, 1)),                # element loads, tiny class
])
def test_noalloc_equals_atb_on_consecutive_calls(first, second):
    for i, (dt, n, k, m) in enumerate((first, second)):
        C0, C1 = both(dt, n, k, m, 2 * i + 1)
        assert torch.equal(C0, C1), f"call {i}: {dt} n={n} k={k} m={m}"
    C0, C1 = both(*second, 9, accumulate=True)
    assert torch.equal(C0, C1)


def test_noalloc_argument_errors_and_the_size_limit():
    L = _lib.lib()
    fake = 1 << 20
    assert L.tsm_atb_noalloc(_lib.F64, _lib.F64, -1, 8, 8, fake, 8, fake, 8, fake, 8, 0,
                             None) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_atb_noalloc(_lib.F64, _lib.BF16, 10, 8, 8, fake, 8, fake, 8, fake, 8, 0,
                             None) == _lib.ERR_UNSUPPORTED
    # every default plan on this GPU fits the static array
    for dt in (_lib.F64, _lib.F32, _lib.BF16):
        for k in (1, 2, 4, 8, 16, 32, 64):
            for m in (1, 2, 4, 8, 16, 32, 64):
                assert L.tsm_atb_workspace_size(dt, 1 << 26, k, m, None) <= _lib.NOALLOC_SCRATCH_BYTES


def test_noalloc_inside_a_captured_graph():
    """After the first call on the device, calls can be captured: every
    replay gives the eager bits."""
    L = _lib.lib()
    n, k, m = 300007, 8, 8
    A = ops.random_matrix(n, k, torch.float64, seed=SEED, stream_id=11)
    B = ops.random_matrix(n, m, torch.float64, seed=SEED, stream_id=12)
    C = torch.empty((k, m), dtype=torch.float64, device="cuda")

    def call():
        st = L.tsm_atb_noalloc(_lib.F64, _lib.F64, n, k, m, A.data_ptr(), k, B.data_ptr(), m,
                               C.data_ptr(), m, 0, torch.cuda.current_stream().cuda_stream)
        assert st == _lib.SUCCESS

    call()
    torch.cuda.synchronize()
    eager = C.clone()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        call()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        call()
    for i in range(3):
        C.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, eager), f"replay {i}"
        call()
        torch.cuda.synchronize()
        assert torch.equal(C, eager)
