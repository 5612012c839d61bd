"""PyTorch-ROCm front end of libtsm: the same entry points on torch tensors.

Tensors must live on a HIP device; the current torch stream is passed to the
library, so calls order with surrounding torch work.  ``atb`` is also
registered as the torch operator ``torch.ops.tsm.atb`` and differentiates
through hand-written backward kernels (``tsm_atb_backward``).
"""
from __future__ import annotations

import ctypes as ct

import torch

from . import _lib

_IN_TYPES = {torch.float64: _lib.F64, torch.float32: _lib.F32, torch.bfloat16: _lib.BF16}
_OUT_TYPES = {torch.float64: _lib.F64, torch.float32: _lib.F32}


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _check_tall(X: torch.Tensor, name: str) -> None:
    if X.dim() != 2:
        raise ValueError(f"{name} must be 2-D (rows x width), got shape {tuple(X.shape)}")
    if not X.is_cuda:
        raise ValueError(f"{name} must be a HIP device tensor")
    if X.stride(1) != 1 or (X.shape[0] > 1 and X.stride(0) < X.shape[1]):
        raise ValueError(f"{name} must be row-major with unit column stride")
    if X.dtype not in _IN_TYPES:
        raise TypeError(f"{name}: unsupported dtype {X.dtype}")


def _ld(X: torch.Tensor) -> int:
    return X.stride(0) if X.shape[0] > 1 else X.shape[1]


def workspace(n: int, k: int, m: int | None = None, dtype: torch.dtype = torch.float64, *,
              config: _lib.KernelConfig | None = None, device="cuda") -> torch.Tensor:
    """A zeroed workspace for ``atb`` (``m`` given) or ``ata`` (``m`` None) on
    inputs of n rows, k (and m) columns and ``dtype``: pass it as
    ``workspace=`` to run those calls without the library's per-stream pool,
    for instance inside a captured graph.  Calls that share one must be
    ordered on one stream; it stays valid for any call of at most these
    sizes and the same kernel request."""
    L = _lib.lib()
    cfg = ct.byref(config) if config is not None else None
    if m is None:
        size = L.tsm_ata_workspace_size(_IN_TYPES[dtype], n, k, cfg)
    else:
        size = L.tsm_atb_workspace_size(_IN_TYPES[dtype], n, k, m, cfg)
    if size == 0:
        raise ValueError("invalid shape or kernel request for a workspace")
    return torch.zeros(size, dtype=torch.uint8, device=device)


def _ws(ws: torch.Tensor | None, device):
    if ws is None:
        return None, 0
    if ws.device != device or not ws.is_contiguous():
        raise ValueError("workspace must be a contiguous tensor on the inputs' device")
    return ws.data_ptr(), ws.numel() * ws.element_size()


def atb(A: torch.Tensor, B: torch.Tensor, *, out_dtype: torch.dtype | None = None,
        out: torch.Tensor | None = None, accumulate: bool = False,
        config: _lib.KernelConfig | None = None,
        workspace: torch.Tensor | None = None) -> torch.Tensor:
    """C = A^T B (or C += A^T B with ``accumulate`` and ``out``), no autograd.

    A: n x k, B: n x m, same dtype (float64, float32 or bfloat16).  C is k x m
    in ``out_dtype`` (float64 by default for float64 inputs, float32 otherwise).
    ``workspace`` (see ``ops.workspace``) replaces the library's per-stream
    pool, which allocates on a stream's first call."""
    _check_tall(A, "A")
    _check_tall(B, "B")
    if A.dtype != B.dtype or A.shape[0] != B.shape[0] or A.device != B.device:
        raise ValueError("A and B must share dtype, row count and device")
    n, k = A.shape
    m = B.shape[1]
    if out is None:
        if out_dtype is None:
            out_dtype = torch.float64 if A.dtype == torch.float64 else torch.float32
        if accumulate:
            raise ValueError("accumulate needs an existing `out`")
        out = torch.empty((k, m), dtype=out_dtype, device=A.device)
    if out.dtype not in _OUT_TYPES or out.shape != (k, m) or out.stride(1) != 1:
        raise ValueError("out must be a row-major k x m float64/float32 tensor")
    if out.device != A.device:
        raise ValueError("out must be on the inputs' device")
    L = _lib.lib()
    ws_ptr, ws_bytes = _ws(workspace, A.device)
    with torch.cuda.device(A.device):
        st = L.tsm_atb_ex(_IN_TYPES[A.dtype], _OUT_TYPES[out.dtype], n, k, m,
                          A.data_ptr(), _ld(A), B.data_ptr(), _ld(B), out.data_ptr(),
                          out.stride(0) if k > 1 else m, 1 if accumulate else 0, ws_ptr,
                          ws_bytes, _stream(), ct.byref(config) if config is not None else None)
    _lib.check(st, "tsm_atb")
    return out


def ata(A: torch.Tensor, *, out_dtype: torch.dtype | None = None,
        out: torch.Tensor | None = None, accumulate: bool = False,
        config: _lib.KernelConfig | None = None,
        workspace: torch.Tensor | None = None) -> torch.Tensor:
    """C = A^T A (or C += A^T A with ``accumulate`` and ``out``): the Gram matrix
    of a tall-skinny A, reading A once.  C is exactly symmetric and within the
    error bound of ``atb(A, A)`` (its kernels tile the rows differently, so
    the bits may differ).
    ``workspace`` as for ``atb``."""
    _check_tall(A, "A")
    n, k = A.shape
    if out is None:
        if out_dtype is None:
            out_dtype = torch.float64 if A.dtype == torch.float64 else torch.float32
        if accumulate:
            raise ValueError("accumulate needs an existing `out`")
        out = torch.empty((k, k), dtype=out_dtype, device=A.device)
    if out.dtype not in _OUT_TYPES or out.shape != (k, k) or out.stride(1) != 1:
        raise ValueError("out must be a row-major k x k float64/float32 tensor")
    if out.device != A.device:
        raise ValueError("out must be on the input's device")
    L = _lib.lib()
    ws_ptr, ws_bytes = _ws(workspace, A.device)
    with torch.cuda.device(A.device):
        st = L.tsm_ata_ex(_IN_TYPES[A.dtype], _OUT_TYPES[out.dtype], n, k, A.data_ptr(),
                          _ld(A), out.data_ptr(), out.stride(0) if k > 1 else k,
                          1 if accumulate else 0, ws_ptr, ws_bytes, _stream(),
                          ct.byref(config) if config is not None else None)
    _lib.check(st, "tsm_ata")
    return out


def tsmm(A: torch.Tensor, X: torch.Tensor) -> torch.Tensor:
    """Y = A X for a tall-skinny A (n x k) and a small X (k x m, float64 or
    float32); Y has A's dtype."""
    _check_tall(A, "A")
    n, k = A.shape
    X = X.contiguous()
    if X.dim() != 2 or X.shape[0] != k or X.dtype not in _OUT_TYPES or X.device != A.device:
        raise ValueError("X must be a k x m float64/float32 tensor on A's device")
    m = X.shape[1]
    Y = torch.empty((n, m), dtype=A.dtype, device=A.device)
    L = _lib.lib()
    with torch.cuda.device(A.device):
        st = L.tsm_tsmm(_IN_TYPES[A.dtype], _OUT_TYPES[X.dtype], n, k, m, A.data_ptr(), _ld(A),
                        X.data_ptr(), m, Y.data_ptr(), m, _stream())
    _lib.check(st, "tsm_tsmm")
    return Y


def atb_backward(A: torch.Tensor, B: torch.Tensor, dC: torch.Tensor, *,
                 need_dA: bool = True, need_dB: bool = True):
    """(dA, dB) = (B dC^T, A dC) in the dtype of A and B; either may be skipped."""
    _check_tall(A, "A")
    _check_tall(B, "B")
    if A.dtype != B.dtype or A.shape[0] != B.shape[0] or A.device != B.device:
        raise ValueError("A and B must share dtype, row count and device")
    n, k = A.shape
    m = B.shape[1]
    if not isinstance(dC, torch.Tensor) or dC.device != A.device:
        raise ValueError("dC must be a tensor on the inputs' device")
    dC = dC.contiguous()
    if dC.shape != (k, m) or dC.dtype not in _OUT_TYPES:
        raise ValueError("dC must be k x m float64/float32")
    dA = torch.empty((n, k), dtype=A.dtype, device=A.device) if need_dA else None
    dB = torch.empty((n, m), dtype=B.dtype, device=B.device) if need_dB else None
    L = _lib.lib()
    with torch.cuda.device(A.device):
        st = L.tsm_atb_backward(_IN_TYPES[A.dtype], _OUT_TYPES[dC.dtype], n, k, m,
                                A.data_ptr(), _ld(A), B.data_ptr(), _ld(B), dC.data_ptr(), m,
                                dA.data_ptr() if dA is not None else None, k,
                                dB.data_ptr() if dB is not None else None, m, _stream())
    _lib.check(st, "tsm_atb_backward")
    return dA, dB


def fill_random(X: torch.Tensor, *, seed: int, stream_id: int, dist: int = _lib.DIST_UNIFORM
                ) -> torch.Tensor:
    """Fill a row-major device matrix with the counter-based synthetic inputs."""
    _check_tall(X, "X")
    rows, cols = X.shape
    L = _lib.lib()
    with torch.cuda.device(X.device):
        st = L.tsm_fill_random(_IN_TYPES[X.dtype], dist, seed, stream_id, X.data_ptr(), rows,
                               cols, _ld(X), _stream())
    _lib.check(st, "tsm_fill_random")
    return X


def random_matrix(rows: int, cols: int, dtype: torch.dtype = torch.float64, *, seed: int,
                  stream_id: int, dist: int = _lib.DIST_UNIFORM, device="cuda") -> torch.Tensor:
    X = torch.empty((rows, cols), dtype=dtype, device=device)
    return fill_random(X, seed=seed, stream_id=stream_id, dist=dist)


def _problems(As, Bs, Cs):
    if not (len(As) == len(Bs) == len(Cs)):
        raise ValueError("As, Bs and outs must have the same length")
    arr = (_lib.Problem * max(len(As), 1))()
    dtype = As[0].dtype if As else torch.float64
    for i, (A, B, C) in enumerate(zip(As, Bs, Cs)):
        _check_tall(A, "A")
        _check_tall(B, "B")
        if A.dtype != dtype or B.dtype != dtype or A.shape[0] != B.shape[0]:
            raise ValueError(f"problem {i}: A and B must share the batch dtype and row count")
        if A.device != As[0].device or B.device != A.device or C.device != A.device:
            raise ValueError(f"problem {i}: A, B and out must be on the batch's device")
        k, m = A.shape[1], B.shape[1]
        if C.shape != (k, m) or C.dtype != Cs[0].dtype or C.dtype not in _OUT_TYPES \
                or C.stride(1) != 1:
            raise ValueError(f"problem {i}: out must be a row-major k x m float64/float32 tensor")
        arr[i] = _lib.Problem(A.data_ptr(), B.data_ptr(), C.data_ptr(), A.shape[0], _ld(A), _ld(B),
                              C.stride(0) if k > 1 else m, k, m)
    return arr, dtype


def atb_batched(As, Bs, *, out_dtype: torch.dtype = torch.float64, outs=None,
                accumulate: bool = False):
    """C_i = A_i^T B_i for every pair in ONE persistent launch; returns the C_i."""
    if outs is None:
        if accumulate:
            raise ValueError("accumulate needs existing `outs`")
        outs = [torch.empty((A.shape[1], B.shape[1]), dtype=out_dtype, device=A.device)
                for A, B in zip(As, Bs)]
    if not As:
        return outs
    arr, dtype = _problems(As, Bs, outs)
    L = _lib.lib()
    with torch.cuda.device(As[0].device):
        st = L.tsm_atb_batched(_IN_TYPES[dtype], _OUT_TYPES[outs[0].dtype], len(As), arr,
                               1 if accumulate else 0, None, 0, _stream())
    _lib.check(st, "tsm_atb_batched")
    return outs


class BatchedPlan:
    """A batched problem list uploaded once; each run() is one kernel launch
    (no allocation or copy), so it can be captured in a HIP graph.  The
    tensors must stay alive and in place for the plan's lifetime."""

    def __init__(self, As, Bs, outs, *, accumulate: bool = False):
        self._keep = (list(As), list(Bs), list(outs))
        arr, dtype = _problems(As, Bs, outs)
        self._h = ct.c_void_p()
        L = _lib.lib()
        with torch.cuda.device(As[0].device):
            st = L.tsm_batched_plan_create(_IN_TYPES[dtype], _OUT_TYPES[outs[0].dtype], len(As),
                                           arr, 1 if accumulate else 0, ct.byref(self._h))
        _lib.check(st, "tsm_batched_plan_create")
        items, grid = ct.c_int64(), ct.c_int64()
        _lib.check(L.tsm_batched_plan_info(self._h, ct.byref(items), ct.byref(grid)),
                   "tsm_batched_plan_info")
        self.items, self.grid = items.value, grid.value

    def run(self) -> None:
        _lib.check(_lib.lib().tsm_batched_plan_execute(self._h, _stream()),
                   "tsm_batched_plan_execute")

    def close(self) -> None:
        if self._h:
            _lib.check(_lib.lib().tsm_batched_plan_destroy(self._h), "tsm_batched_plan_destroy")
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _AtbFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B, out_dtype):
        ctx.save_for_backward(A, B)
        return atb(A, B, out_dtype=out_dtype)

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        need_dA, need_dB = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_dA or need_dB):
            return None, None, None
        if dC.dtype not in _OUT_TYPES:
            dC = dC.float()
        dA, dB = atb_backward(A, B, dC, need_dA=need_dA, need_dB=need_dB)
        return dA, dB, None


def atb_autograd(A: torch.Tensor, B: torch.Tensor, out_dtype: torch.dtype | None = None):
    """Differentiable C = A^T B (gradients from the hand-written backward kernels)."""
    if out_dtype is None:
        out_dtype = torch.float64 if A.dtype == torch.float64 else torch.float32
    return _AtbFunction.apply(A, B, out_dtype)


# ---- torch operator registration ---------------------------------------------------
@torch.library.custom_op("tsm::atb", mutates_args=())
def _tsm_atb_op(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    return atb(A, B)


@_tsm_atb_op.register_fake
def _(A, B):
    out = torch.float64 if A.dtype == torch.float64 else torch.float32
    return A.new_empty((A.shape[1], B.shape[1]), dtype=out)


def _op_setup_context(ctx, inputs, output):
    ctx.save_for_backward(*inputs)


def _op_backward(ctx, dC):
    A, B = ctx.saved_tensors
    if dC.dtype not in _OUT_TYPES:
        dC = dC.float()
    return atb_backward(A, B, dC, need_dA=ctx.needs_input_grad[0],
                        need_dB=ctx.needs_input_grad[1])


_tsm_atb_op.register_autograd(_op_backward, setup_context=_op_setup_context)
