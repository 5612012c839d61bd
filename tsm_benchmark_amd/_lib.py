"""ctypes binding of libtsm.so (the C ABI declared in include/tsm/tsm.h).

There is no fallback: if the shared library is missing or fails to load,
every entry point raises, naming the file and the build command.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtsm.so")

# tsm_dtype
F64, F32, BF16 = 0, 1, 2
# tsm_dist
DIST_UNIFORM, DIST_POW2 = 0, 1
# tsm_status
SUCCESS, ERR_INVALID_ARGUMENT, ERR_UNSUPPORTED, ERR_MISALIGNED, ERR_WORKSPACE, ERR_HIP, \
    ERR_NO_KERNEL = range(7)

# Every function include/tsm/tsm.h declares (checked by tests/test_capi.py).
EXPORTS = (
    "tsm_atb", "tsm_atb_ex", "tsm_atb_workspace_size", "tsm_atb_backward",
    "tsm_atb_batched_workspace_size", "tsm_atb_batched", "tsm_fill_random",
    "tsm_strerror", "tsm_last_hip_error", "tsm_version", "tsm_num_kernels",
    "tsm_kernel_info", "tsm_atb_plan", "tsm_batched_plan_create", "tsm_batched_plan_execute",
    "tsm_batched_plan_info", "tsm_batched_plan_destroy", "tsm_ata", "tsm_ata_ex", "tsm_ata_workspace_size",
    "tsm_tsmm", "tsm_atb_noalloc",
)
# include/tsm/tsm.h TSM_NOALLOC_SCRATCH_BYTES: the static workspace of tsm_atb_noalloc
NOALLOC_SCRATCH_BYTES = 256 + (16 << 20)


class KernelConfig(ct.Structure):
    """tsm_kernel_config: zero fields take the tuned default."""
    _fields_ = [("tm", ct.c_int), ("tn", ct.c_int), ("rows", ct.c_int), ("pf", ct.c_int),
                ("wgs_per_cu", ct.c_int), ("grid", ct.c_int64), ("load_policy", ct.c_int),
                ("tile_order", ct.c_int), ("engine", ct.c_int)]


class KernelDesc(ct.Structure):
    _fields_ = [("name", ct.c_char_p), ("in_type", ct.c_int), ("kc", ct.c_int),
                ("mc", ct.c_int), ("tm", ct.c_int), ("tn", ct.c_int), ("rows", ct.c_int),
                ("pf", ct.c_int), ("load_policy", ct.c_int), ("lds_bytes", ct.c_int),
                ("engine", ct.c_int)]


class Problem(ct.Structure):
    """tsm_problem: one entry of a batched call."""
    _fields_ = [("A", ct.c_void_p), ("B", ct.c_void_p), ("C", ct.c_void_p),
                ("n", ct.c_int64), ("lda", ct.c_int64), ("ldb", ct.c_int64),
                ("ldc", ct.c_int64), ("k", ct.c_int), ("m", ct.c_int)]


class TsmError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        msg = f"{what}: {strerror(status)} (status {status})"
        if status == ERR_HIP:
            msg += f", hipError {lib().tsm_last_hip_error()}"
        super().__init__(msg)


_lib = None


def lib() -> ct.CDLL:
    """Load libtsm.so once; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m tsm_benchmark_amd.build` "
            "(there is no CPU fallback)")
    # torch's wheel carries its own libamdhip64 (SONAME libamdhip64.so.7, but
    # torch's libraries NEED it as "libamdhip64.so").  Loading torch first makes
    # libtsm bind to that same runtime; loading libtsm first would put a second
    # HIP runtime into the process and torch's stream handles would be foreign
    # to it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ct.CDLL(LIB_PATH)
    P, I, I64, SZ, U64 = ct.c_void_p, ct.c_int, ct.c_int64, ct.c_size_t, ct.c_uint64
    L.tsm_atb.restype = I
    L.tsm_atb.argtypes = [I, I, I64, I, I, P, I64, P, I64, P, I64, I, P, SZ, P]
    L.tsm_atb_ex.restype = I
    L.tsm_atb_ex.argtypes = [I, I, I64, I, I, P, I64, P, I64, P, I64, I, P, SZ, P,
                             ct.POINTER(KernelConfig)]
    L.tsm_atb_noalloc.restype = I
    L.tsm_atb_noalloc.argtypes = [I, I, I64, I, I, P, I64, P, I64, P, I64, I, P]
    L.tsm_atb_workspace_size.restype = SZ
    L.tsm_atb_workspace_size.argtypes = [I, I64, I, I, ct.POINTER(KernelConfig)]
    L.tsm_atb_backward.restype = I
    L.tsm_atb_backward.argtypes = [I, I, I64, I, I, P, I64, P, I64, P, I64, P, I64, P, I64, P]
    L.tsm_ata.restype = I
    L.tsm_ata.argtypes = [I, I, I64, I, P, I64, P, I64, I, P, SZ, P]
    L.tsm_ata_ex.restype = I
    L.tsm_ata_ex.argtypes = [I, I, I64, I, P, I64, P, I64, I, P, SZ, P, ct.POINTER(KernelConfig)]
    L.tsm_ata_workspace_size.restype = SZ
    L.tsm_ata_workspace_size.argtypes = [I, I64, I, ct.POINTER(KernelConfig)]
    L.tsm_tsmm.restype = I
    L.tsm_tsmm.argtypes = [I, I, I64, I, I, P, I64, P, I64, P, I64, P]
    L.tsm_atb_batched_workspace_size.restype = SZ
    L.tsm_atb_batched_workspace_size.argtypes = [I, I, ct.POINTER(Problem)]
    L.tsm_atb_batched.restype = I
    L.tsm_atb_batched.argtypes = [I, I, I, ct.POINTER(Problem), I, P, SZ, P]
    L.tsm_fill_random.restype = I
    L.tsm_fill_random.argtypes = [I, I, U64, U64, P, I64, I64, I64, P]
    L.tsm_strerror.restype = ct.c_char_p
    L.tsm_strerror.argtypes = [I]
    L.tsm_last_hip_error.restype = I
    L.tsm_version.restype = I
    L.tsm_num_kernels.restype = I
    L.tsm_kernel_info.restype = I
    L.tsm_kernel_info.argtypes = [I, ct.POINTER(KernelDesc)]
    L.tsm_batched_plan_create.restype = I
    L.tsm_batched_plan_create.argtypes = [I, I, I, ct.POINTER(Problem), I, ct.POINTER(P)]
    L.tsm_batched_plan_execute.restype = I
    L.tsm_batched_plan_execute.argtypes = [P, P]
    L.tsm_batched_plan_info.restype = I
    L.tsm_batched_plan_info.argtypes = [P, ct.POINTER(I64), ct.POINTER(I64)]
    L.tsm_batched_plan_destroy.restype = I
    L.tsm_batched_plan_destroy.argtypes = [P]
    L.tsm_atb_plan.restype = I
    L.tsm_atb_plan.argtypes = [I, I64, I, I, ct.POINTER(KernelConfig), ct.POINTER(I),
                               ct.POINTER(I64)]
    _lib = L
    return L


def strerror(status: int) -> str:
    return lib().tsm_strerror(status).decode()


def check(status: int, what: str) -> None:
    if status != SUCCESS:
        raise TsmError(status, what)


def kernels() -> list[dict]:
    """The generated kernel table (name and configuration of every kernel)."""
    L = lib()
    out = []
    for i in range(L.tsm_num_kernels()):
        d = KernelDesc()
        check(L.tsm_kernel_info(i, ct.byref(d)), "tsm_kernel_info")
        out.append({"index": i, "name": d.name.decode(), "in_type": d.in_type, "kc": d.kc,
                    "mc": d.mc, "tm": d.tm, "tn": d.tn, "rows": d.rows, "pf": d.pf,
                    "nt": 1 if d.load_policy == 2 else 0, "lds_bytes": d.lds_bytes,
                    "engine": d.engine})
    return out


ENGINE_VALU, ENGINE_MFMA_LDS, ENGINE_DIRECT = 1, 2, 3


def make_config(tm=0, tn=0, rows=0, pf=0, wgs_per_cu=0, grid=0, nt=None,
                interleave=None, mfma=None, engine=None) -> KernelConfig:
    """nt: None = tuned default, 0 = cached loads, 1 = non-temporal loads;
    interleave: None = tuned default, 0 = contiguous tile runs, 1 = interleaved;
    mfma: None = tuned default, False = VALU register tiles, True = matrix cores
    through LDS; engine (overrides mfma): 1 VALU, 2 matrix cores through LDS,
    3 matrix cores fed from global memory (pf = 4-row steps in flight)."""
    policy = 0 if nt is None else (2 if nt else 1)
    order = 0 if interleave is None else (2 if interleave else 1)
    if engine is None:
        engine = 0 if mfma is None else (ENGINE_MFMA_LDS if mfma else ENGINE_VALU)
    return KernelConfig(tm, tn, rows, pf, wgs_per_cu, grid, policy, order, engine)


def config_of(kd: dict, **kw) -> KernelConfig:
    """Config that selects exactly the kernel kd (an entry of kernels())."""
    return make_config(tm=kd["tm"], tn=kd["tn"], rows=kd["rows"], pf=kd["pf"], nt=kd["nt"],
                       engine=kd["engine"], **kw)


def plan(in_type: int, n: int, k: int, m: int, config: KernelConfig | None = None):
    """(kernel index, grid) that tsm_atb would use; needs a GPU for the CU count."""
    idx, grid = ct.c_int(-1), ct.c_int64(0)
    check(lib().tsm_atb_plan(in_type, n, k, m, ct.byref(config) if config else None,
                             ct.byref(idx), ct.byref(grid)), "tsm_atb_plan")
    return idx.value, grid.value
