"""Argument checks of the torch front end and of the C ABI around graph
capture: every case is rejected before a kernel is launched."""
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import _lib, ops  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    torch.cuda.set_device(0)
    yield


def dev(n, w, dt=torch.float64):
    return torch.zeros((n, w), dtype=dt, device="cuda")


def test_front_end_rejects_mismatched_operands():
    A, B = dev(100, 4), dev(100, 3)
    with pytest.raises(ValueError):
        ops.atb(A, dev(99, 3))                       # row counts differ
    with pytest.raises(ValueError):
        ops.atb(A, dev(100, 3, torch.float32))       # dtypes differ
    with pytest.raises(ValueError):
        ops.atb(A, B, out=torch.zeros((4, 3), dtype=torch.float64))   # out on the host
    with pytest.raises(ValueError):
        ops.ata(A, out=torch.zeros((4, 4), dtype=torch.float64))
    G = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        ops.atb_backward(A, dev(99, 3), G)           # B shorter than A
    with pytest.raises(ValueError):
        ops.atb_backward(A, dev(100, 3, torch.float32), G)
    with pytest.raises(ValueError):
        ops.atb_backward(A, B, G.cpu())              # dC on the host
    with pytest.raises(ValueError):
        ops.atb_batched([A], [B], outs=[torch.zeros((4, 3), dtype=torch.float64)])
    with pytest.raises(ValueError):
        ops.BatchedPlan([A], [B], [torch.zeros((4, 3), dtype=torch.float64)])


def test_capture_without_a_workspace_is_refused():
    """The per-stream pool cannot serve a captured call: the call returns
    TSM_ERR_WORKSPACE before it allocates or launches anything, and a caller
    workspace makes the same capture work."""
    n, k, m = 4096, 8, 8
    A, B = dev(n, k) + 1.0, dev(n, m) + 2.0
    C = torch.empty((k, m), dtype=torch.float64, device="cuda")
    ws = ops.workspace(n, k, m)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ops.atb(A, B, out=C, workspace=ws)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    status = {}
    with torch.cuda.graph(g):
        try:
            ops.atb(A, B, out=C)
        except _lib.TsmError as e:
            status["pool"] = e.status
        try:
            ops.atb_batched([A], [B], outs=[C])
        except _lib.TsmError as e:
            status["batched"] = e.status
        ops.atb(A, B, out=C, workspace=ws)
    assert status == {"pool": _lib.ERR_WORKSPACE, "batched": _lib.ERR_WORKSPACE}
    C.zero_()
    g.replay()
    torch.cuda.synchronize()
    assert torch.all(C == 2.0 * n)


def test_batched_call_with_a_workspace_is_refused_under_capture():
    """tsm_atb_batched uploads its problem list through a reused host buffer,
    which a graph cannot replay: under capture it returns TSM_ERR_UNSUPPORTED
    (a plan is the graph-capturable form)."""
    n = 1000
    A, B = dev(n, 4) + 1.0, dev(n, 2) + 3.0
    C = torch.empty((4, 2), dtype=torch.float64, device="cuda")
    L = _lib.lib()
    P = (_lib.Problem * 1)()
    P[0] = _lib.Problem(A.data_ptr(), B.data_ptr(), C.data_ptr(), n, 4, 2, 2, 4, 2)
    wsb = L.tsm_atb_batched_workspace_size(_lib.F64, 1, P)
    ws = torch.zeros(wsb, dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    assert L.tsm_atb_batched(_lib.F64, _lib.F64, 1, P, 0, ws.data_ptr(), wsb,
                             torch.cuda.current_stream().cuda_stream) == _lib.SUCCESS
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        st = L.tsm_atb_batched(_lib.F64, _lib.F64, 1, P, 0, ws.data_ptr(), wsb,
                               torch.cuda.current_stream().cuda_stream)
        C.zero_()
    assert st == _lib.ERR_UNSUPPORTED
    g.replay()
    torch.cuda.synchronize()
    # outside capture the same call runs
    assert L.tsm_atb_batched(_lib.F64, _lib.F64, 1, P, 0, ws.data_ptr(), wsb,
                             torch.cuda.current_stream().cuda_stream) == _lib.SUCCESS
    torch.cuda.synchronize()
    assert torch.all(C == 3.0 * n)


def test_plan_destroy_waits_for_its_launches_on_every_stream():
    n = 1 << 16
    As = [dev(n, 8) + 1.0, dev(n, 4) + 0.5]
    Bs = [dev(n, 8) + 2.0, dev(n, 2) + 4.0]
    outs = [torch.empty((8, 8), dtype=torch.float64, device="cuda"),
            torch.empty((4, 2), dtype=torch.float64, device="cuda")]
    plan = ops.BatchedPlan(As, Bs, outs)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s1):
        plan.run()
    s2.wait_stream(s1)  # launches of one plan share its workspace: keep them ordered
    with torch.cuda.stream(s2):
        plan.run()
    plan.close()  # waits for both launches, then frees the plan's buffer
    assert torch.all(outs[0] == 2.0 * n) and torch.all(outs[1] == 2.0 * n)
