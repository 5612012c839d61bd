"""The C-ABI library without a GPU: it loads, exports every function that
include/tsm/tsm.h declares, and rejects invalid arguments with the documented
status codes before any HIP call (so these paths behave the same on the CPU).
"""
import ctypes as ct
import os
import re

import pytest

from tsm_benchmark_amd import _lib, codegen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tsm", "tsm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsm_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert L.tsm_version() == 100


def test_strerror_covers_every_status():
    for s in range(7):
        assert _lib.strerror(s) and _lib.strerror(s) != "unknown status"
    assert _lib.strerror(99) == "unknown status"


def test_kernel_table_matches_generator():
    cfgs, defaults = codegen.library_kernels(codegen.load_tuning())
    table = _lib.kernels()
    assert [k["name"] for k in table] == [c.name for c in cfgs]
    for k, c in zip(table, cfgs):
        assert (k["kc"], k["mc"], k["tm"], k["tn"], k["rows"], k["pf"], k["nt"], k["lds_bytes"],
                k["engine"]) == (c.kc, c.mc, c.tm, c.tn, c.rows, c.pf, c.nt, c.lds_bytes, c.engine)
    # every (dtype, k class, m class) has a kernel
    have = {(k["in_type"], k["kc"], k["mc"]) for k in table}
    for dt in range(3):
        for kc in codegen.CLASSES:
            for mc in codegen.CLASSES:
                assert (dt, kc, mc) in have


FAKE = 0x10000  # never dereferenced: validation fails first


def atb(**kw):
    a = dict(in_type=_lib.F64, out_type=_lib.F64, n=10, k=8, m=8, A=FAKE, lda=8, B=FAKE, ldb=8,
             C=FAKE, ldc=8, acc=0, ws=None, wsb=0, stream=None)
    a.update(kw)
    return _lib.lib().tsm_atb(a["in_type"], a["out_type"], a["n"], a["k"], a["m"], a["A"],
                              a["lda"], a["B"], a["ldb"], a["C"], a["ldc"], a["acc"], a["ws"],
                              a["wsb"], a["stream"])


@pytest.mark.parametrize("kw,status", [
    (dict(n=-1), _lib.ERR_INVALID_ARGUMENT),
    (dict(k=0), _lib.ERR_INVALID_ARGUMENT),
    (dict(m=65), _lib.ERR_INVALID_ARGUMENT),
    (dict(lda=7), _lib.ERR_INVALID_ARGUMENT),
    (dict(ldb=4), _lib.ERR_INVALID_ARGUMENT),
    (dict(ldc=7), _lib.ERR_INVALID_ARGUMENT),
    (dict(A=None), _lib.ERR_INVALID_ARGUMENT),
    (dict(C=None), _lib.ERR_INVALID_ARGUMENT),
    (dict(A=FAKE + 4), _lib.ERR_MISALIGNED),
    (dict(in_type=_lib.F32, B=FAKE + 2), _lib.ERR_MISALIGNED),
    (dict(out_type=_lib.F64, C=FAKE + 4), _lib.ERR_MISALIGNED),
    (dict(in_type=7), _lib.ERR_UNSUPPORTED),
    (dict(out_type=_lib.BF16), _lib.ERR_UNSUPPORTED),
    (dict(n=1 << 62), _lib.ERR_INVALID_ARGUMENT),
])
def test_atb_argument_errors(kw, status):
    assert atb(**kw) == status


def test_atb_config_errors():
    L = _lib.lib()
    for bad in (_lib.make_config(pf=5), _lib.make_config(pf=-1), _lib.make_config(engine=4)):
        st = L.tsm_atb_ex(_lib.F64, _lib.F64, 10, 8, 8, FAKE, 8, FAKE, 8, FAKE, 8, 0, None, 0, None,
                          ct.byref(bad))
        assert st == _lib.ERR_INVALID_ARGUMENT
    # a valid request no kernel satisfies: the direct engine needs widths of 16 or more
    direct = _lib.make_config(engine=_lib.ENGINE_DIRECT)
    st = L.tsm_atb_ex(_lib.F64, _lib.F64, 10, 8, 8, FAKE, 8, FAKE, 8, FAKE, 8, 0, None, 0,
                      None, ct.byref(direct))
    assert st == _lib.ERR_NO_KERNEL


def test_backward_argument_errors():
    L = _lib.lib()

    def bwd(in_type=_lib.F64, g=_lib.F64, n=10, k=4, m=4, A=FAKE, lda=4, B=FAKE, ldb=4, dC=FAKE,
            lddc=4, dA=FAKE, ldda=4, dB=FAKE, lddb=4):
        return L.tsm_atb_backward(in_type, g, n, k, m, A, lda, B, ldb, dC, lddc, dA, ldda, dB,
                                  lddb, None)

    assert bwd(k=0) == _lib.ERR_INVALID_ARGUMENT
    assert bwd(dC=None) == _lib.ERR_INVALID_ARGUMENT
    assert bwd(ldda=3) == _lib.ERR_INVALID_ARGUMENT
    assert bwd(dB=FAKE + 2) == _lib.ERR_MISALIGNED
    assert bwd(g=_lib.BF16) == _lib.ERR_UNSUPPORTED
    assert bwd(B=None) == _lib.ERR_INVALID_ARGUMENT
    # (n - 1) * ld would leave int64: refused before any launch
    assert bwd(n=1 << 60, lda=64, k=4) == _lib.ERR_INVALID_ARGUMENT
    assert bwd(n=1 << 59, lddb=64) == _lib.ERR_INVALID_ARGUMENT
    # nothing to do: succeeds without touching the device
    assert bwd(dA=None, dB=None) == _lib.SUCCESS
    assert bwd(n=0) == _lib.SUCCESS


def test_ata_argument_errors():
    L = _lib.lib()

    def ata(in_type=_lib.F64, out_type=_lib.F64, n=10, k=8, A=FAKE, lda=8, C=FAKE, ldc=8):
        return L.tsm_ata(in_type, out_type, n, k, A, lda, C, ldc, 0, None, 0, None)

    assert ata(k=0) == _lib.ERR_INVALID_ARGUMENT
    assert ata(k=65) == _lib.ERR_INVALID_ARGUMENT
    assert ata(lda=7) == _lib.ERR_INVALID_ARGUMENT
    assert ata(ldc=7) == _lib.ERR_INVALID_ARGUMENT
    assert ata(A=None) == _lib.ERR_INVALID_ARGUMENT
    assert ata(C=None) == _lib.ERR_INVALID_ARGUMENT
    assert ata(A=FAKE + 4) == _lib.ERR_MISALIGNED
    assert ata(in_type=9) == _lib.ERR_UNSUPPORTED
    assert ata(out_type=_lib.BF16) == _lib.ERR_UNSUPPORTED
    assert L.tsm_ata_workspace_size(_lib.F64, -1, 8, None) == 0
    assert L.tsm_ata_workspace_size(_lib.F64, 10, 65, None) == 0
    bad = _lib.make_config(pf=5)
    assert L.tsm_ata_ex(_lib.F64, _lib.F64, 10, 8, FAKE, 8, FAKE, 8, 0, None, 0, None,
                        ct.byref(bad)) == _lib.ERR_INVALID_ARGUMENT


def test_tsmm_argument_errors():
    L = _lib.lib()

    def tsmm(in_type=_lib.F64, x_type=_lib.F64, n=10, k=4, m=8, A=FAKE, lda=4, X=FAKE, ldx=8,
             Y=FAKE, ldy=8):
        return L.tsm_tsmm(in_type, x_type, n, k, m, A, lda, X, ldx, Y, ldy, None)

    assert tsmm(k=0) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(m=65) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(lda=3) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(ldx=7) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(ldy=7) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(X=None) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(A=None) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(Y=None) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(Y=FAKE + 4) == _lib.ERR_MISALIGNED
    assert tsmm(x_type=_lib.BF16) == _lib.ERR_UNSUPPORTED
    assert tsmm(n=1 << 60, ldy=64) == _lib.ERR_INVALID_ARGUMENT
    assert tsmm(n=0) == _lib.SUCCESS


def test_fill_argument_errors():
    L = _lib.lib()
    assert L.tsm_fill_random(_lib.F64, 5, 1, 1, FAKE, 4, 4, 4, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_fill_random(_lib.F64, 0, 1, 1, FAKE, 4, 4, 3, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_fill_random(_lib.F32, 0, 1, 1, FAKE + 2, 4, 4, 4, None) == _lib.ERR_MISALIGNED
    assert L.tsm_fill_random(9, 0, 1, 1, FAKE, 4, 4, 4, None) == _lib.ERR_UNSUPPORTED
    assert L.tsm_fill_random(_lib.F64, 0, 1, 1, FAKE, 1 << 60, 8, 8, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_fill_random(_lib.F64, 0, 1, 1, None, 0, 4, 4, None) == _lib.SUCCESS


def test_kernel_info_bounds():
    L = _lib.lib()
    d = _lib.KernelDesc()
    assert L.tsm_kernel_info(-1, ct.byref(d)) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_kernel_info(L.tsm_num_kernels(), ct.byref(d)) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_kernel_info(0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_atb_workspace_size(_lib.F64, -1, 8, 8, None) == 0
    # a kernel request that tsm_atb_ex would reject has no workspace size either
    for bad in (_lib.make_config(pf=5), _lib.make_config(engine=9), _lib.make_config(rows=-1)):
        assert L.tsm_atb_workspace_size(_lib.F64, 10, 8, 8, ct.byref(bad)) == 0
        assert L.tsm_ata_workspace_size(_lib.F64, 10, 8, ct.byref(bad)) == 0


def test_batched_argument_errors():
    L = _lib.lib()
    P = (_lib.Problem * 2)()
    P[0] = _lib.Problem(FAKE, FAKE, FAKE, 10, 8, 8, 8, 8, 8)
    P[1] = _lib.Problem(FAKE, FAKE, FAKE, 10, 8, 8, 8, 8, 0)  # m = 0
    assert L.tsm_atb_batched(_lib.F64, _lib.F64, 2, P, 0, None, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_atb_batched(_lib.F64, _lib.F64, -1, P, 0, None, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_atb_batched(_lib.F64, _lib.BF16, 1, P, 0, None, 0, None) == _lib.ERR_UNSUPPORTED
    # FP32/BF16 problems wider than 32 run as column blocks: the layout of a
    # 64 x 8 problem is that of its two 32 x 8 blocks (no device query)
    W1 = (_lib.Problem * 1)()
    W1[0] = _lib.Problem(FAKE, FAKE, FAKE, 10000, 64, 8, 8, 64, 8)
    W2 = (_lib.Problem * 2)()
    W2[0] = _lib.Problem(FAKE, FAKE, FAKE, 10000, 64, 8, 8, 32, 8)
    W2[1] = _lib.Problem(FAKE + 128, FAKE, FAKE + 32 * 8 * 8, 10000, 64, 8, 8, 32, 8)
    for dt in (_lib.F32, _lib.BF16):
        assert L.tsm_atb_batched_workspace_size(dt, 1, W1) == L.tsm_atb_batched_workspace_size(dt, 2, W2) > 0
    # FP64 has a 64-wide class: no split, another layout
    assert L.tsm_atb_batched_workspace_size(_lib.F64, 1, W1) != L.tsm_atb_batched_workspace_size(_lib.F64, 2, W2)
    assert L.tsm_atb_batched(_lib.F64, _lib.F64, 0, None, 0, None, 0, None) == _lib.SUCCESS
    # the workspace size is a pure function of the problem list (no device query)
    P[1] = _lib.Problem(FAKE, FAKE, FAKE, 10000, 8, 8, 8, 8, 8)
    assert L.tsm_atb_batched_workspace_size(_lib.F64, 2, P) > 0
    assert L.tsm_atb_batched_workspace_size(_lib.F64, 2, P) == L.tsm_atb_batched_workspace_size(_lib.F64, 2, P)
    small = L.tsm_atb_batched_workspace_size(_lib.F64, 2, P) - 8
    assert L.tsm_atb_batched(_lib.F64, _lib.F64, 2, P, 0, FAKE, small, None) == _lib.ERR_WORKSPACE
    assert L.tsm_batched_plan_execute(None, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.tsm_batched_plan_destroy(None) == _lib.ERR_INVALID_ARGUMENT
