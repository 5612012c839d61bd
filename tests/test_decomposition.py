"""Work-decomposition coverage of the restated reference kernels.

Black-box visit counting: with A = 1 and B[r, j] = 4^r (r < 25) every entry of
C equals sum_r count(r, i, j) * 4^r exactly, so its base-4 digits are the
number of times each row reached output (i, j).  With A[r, i] = 4^i and B = 1
(and the mirror) the digits tell which column of A (B) fed output (i, j).
Every configuration the reference's main.cc and generator use (scaled down in
the number of rows) must visit each (row, i, j) exactly once; the sub-group
size hazard of the generated _sg kernels (sycl_codegen_new.py:250-253, 282-287)
is shown as a double count.
"""
import numpy as np
import pytest

NROWS = 25  # 4^24 * 3 < 2^53: all counts stay exact


def _digits(x, ndig):
    x = int(round(x))
    out = []
    for _ in range(ndig):
        out.append(x % 4)
        x //= 4
    assert x == 0, "count overflowed its base-4 digit"
    return out


def row_counts(orc, cfg, k, m):
    A = np.ones((NROWS, k))
    B = np.repeat((4.0 ** np.arange(NROWS))[:, None], m, axis=1)
    C = orc.ref_run(cfg, A, B)
    return np.array([[_digits(C[i, j], NROWS) for j in range(m)] for i in range(k)])


def column_sources(orc, cfg, k, m):
    A = np.repeat((4.0 ** np.arange(k))[None, :], NROWS, axis=0)
    B = np.ones((NROWS, m))
    CA = orc.ref_run(cfg, A, B) / NROWS
    A2 = np.ones((NROWS, k))
    B2 = np.repeat((4.0 ** np.arange(m))[None, :], NROWS, axis=0)
    CB = orc.ref_run(cfg, A2, B2) / NROWS
    return CA, CB


def assert_exact_cover(orc, cfg, k, m):
    cnt = row_counts(orc, cfg, k, m)
    assert np.all(cnt == 1), f"rows visited {cnt.min()}..{cnt.max()} times"
    CA, CB = column_sources(orc, cfg, k, m)
    np.testing.assert_array_equal(CA, np.repeat((4.0 ** np.arange(k))[:, None], m, axis=1))
    np.testing.assert_array_equal(CB, np.repeat((4.0 ** np.arange(m))[None, :], k, axis=0))


GEN_CASES = [  # (M, N, TM, TN, stride, local, sg, unroll)
    (4, 4, 2, 2, 256, 256, 32, 1), (8, 8, 4, 4, 512, 256, 32, 1), (8, 8, 2, 2, 2048, 256, 32, 3),
    (16, 16, 4, 4, 4096, 256, 32, 2), (4, 32, 4, 8, 1024, 256, 32, 1),
    (32, 4, 8, 4, 1024, 256, 32, 4), (1, 1, 1, 1, 64, 64, 32, 1), (8, 8, 8, 8, 64, 64, 64, 2),
]


@pytest.mark.parametrize("case", GEN_CASES)
@pytest.mark.parametrize("variant", ["basic", "lf", "sg", "sg_lf"])
def test_generated_kernels_cover_each_row_once(orc, case, variant):
    M, N, TM, TN, stride, local, sg, unroll = case
    vid = {"basic": orc.REF_GEN_BASIC, "lf": orc.REF_GEN_LF, "sg": orc.REF_GEN_SG,
           "sg_lf": orc.REF_GEN_SG_LF}[variant]
    cfg = orc.ref_cfg(vid, tm=TM, tn=TN, stride=stride, local_size=local, sg_size_gen=sg,
                      unroll=unroll)
    assert_exact_cover(orc, cfg, M, N)


@pytest.mark.parametrize("M,stride,tile", [(4, 4096, 2), (4, 4096, 4), (8, 8192, 2),
                                           (8, 16384, 4), (16, 32768, 4), (2, 64, 4)])
@pytest.mark.parametrize("variant", ["V1", "V2", "V3", "V4", "V5", "V7"])
def test_range_variants_cover_each_row_once(orc, M, stride, tile, variant):
    # main.cc's (M, stride, tile) sets (main.cc:171-192), fewer rows
    cfg = orc.ref_cfg(getattr(orc, "REF_" + variant), tm=tile, stride=stride)
    assert_exact_cover(orc, cfg, M, M)


@pytest.mark.parametrize("M,stride,tile", [(4, 4096, 2), (8, 8192, 4), (16, 32768, 4)])
@pytest.mark.parametrize("variant", ["V6", "V9"])
def test_ndrange_variants_cover_each_row_once(orc, M, stride, tile, variant):
    cfg = orc.ref_cfg(getattr(orc, "REF_" + variant), tm=tile, stride=stride, local_size=32)
    assert_exact_cover(orc, cfg, M, M)


@pytest.mark.parametrize("M", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("num_cu", [1, 4, 256])
def test_adaptive_variant8_covers_each_row_once(orc, M, num_cu):
    cfg = orc.ref_cfg(orc.REF_V8, worker_mult=64, num_cu=num_cu)
    assert_exact_cover(orc, cfg, M, M)


def test_subgroup_size_mismatch_double_counts(orc):
    # The generator asserts stride/sg_gen % num_tiles == 0 with its own
    # sg_size (32), but the kernel's index math uses the runtime sub-group
    # size (64): with stride 1536 there are 24 runtime sub-groups for 16 tiles,
    # k_stride = (24 / 16) * 64 = 64, and sub-groups 16..23 revisit rows 64+ of
    # tiles 0..7.  With A = B = 1 every C entry counts its visits.
    n = 128
    cfg = orc.ref_cfg(orc.REF_GEN_SG, tm=2, stride=1536, local_size=256, sg_size_gen=32,
                      sg_size_rt=64)
    C = orc.ref_run(cfg, np.ones((n, 8)), np.ones((n, 8)))
    # tile p = midx + 4*nidx owns C[a*4 + midx, b*4 + nidx]; tiles 0..7 are nidx 0, 1
    tile = (np.arange(8)[:, None] % 4) + 4 * (np.arange(8)[None, :] % 4)
    np.testing.assert_array_equal(C, np.where(tile < 8, n + 64, n))
    # with the sizes agreeing, the same stride covers every row once
    ok = orc.ref_cfg(orc.REF_GEN_SG, tm=2, stride=1536, local_size=256, sg_size_gen=32,
                     sg_size_rt=32)
    np.testing.assert_array_equal(orc.ref_run(ok, np.ones((n, 8)), np.ones((n, 8))), n)


def test_zero_row_stride_is_refused(orc):
    cfg = orc.ref_cfg(orc.REF_GEN_SG, tm=2, stride=512, local_size=256, sg_size_gen=32,
                      sg_size_rt=64)
    with pytest.raises(ValueError):
        orc.ref_run(cfg, np.ones((4, 8)), np.ones((4, 8)))


def test_restated_kernels_match_exact_oracle(orc):
    n = 1 << 16
    A = orc.gen_matrix(n, 8, seed=1, stream=1)
    B = orc.gen_matrix(n, 8, seed=1, stream=2)
    ex, absC = orc.atb_dot2(A, B, with_abs=True)
    for vid in (orc.REF_GEN_BASIC, orc.REF_GEN_LF, orc.REF_GEN_SG, orc.REF_GEN_SG_LF,
                orc.REF_V4, orc.REF_V6, orc.REF_V9):
        C = orc.ref_run(orc.ref_cfg(vid, tm=4, stride=16384, local_size=32), A, B)
        assert np.all(np.abs(C - ex) <= 4 * np.sqrt(n) * 2.0**-52 * absC + n * 2.0**-53 * absC)
        # and the reference's own acceptance rule (main.cc:29-41)
        assert orc.first_mismatch(ex, C) == -1


def test_stride_rule(orc):
    # default flags of the generator: M in {4, 8, 16}, TM in {2, 4}, L 256, sg 32
    assert orc.suggest_stride([4, 8, 16], [2, 4], 256, 32, 2048) == 524288
    assert orc.validate_stride([4, 8, 16], [2, 4], 524288, 256, 32)
    assert not orc.validate_stride([4, 8, 16], [2, 4], 524288 + 256, 256, 32)
    from tsm_benchmark_amd import codegen

    for Ms, tiles, L, sg in (([4, 8, 16], [2, 4], 256, 32), ([8], [4], 256, 64),
                             ([16, 32], [4, 8], 256, 64)):
        assert codegen.suggest_stride(Ms, tiles, L, sg) == orc.suggest_stride(Ms, tiles, L, sg, 2048)
        s = codegen.suggest_stride(Ms, tiles, L, sg)
        assert (not codegen.validate_stride(Ms, tiles, s, L, sg)) == orc.validate_stride(Ms, tiles, s, L, sg)
