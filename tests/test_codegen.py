"""The kernel generator: configuration rules, CLI validation (the reference
generator's flag checks, sycl_codegen_new.py:606-620), the library kernel set
and the reference-name layer."""
import subprocess
import sys

import pytest

from tsm_benchmark_amd import codegen
from tsm_benchmark_amd.codegen import KernelConfig


def run_cli(*args):
    return subprocess.run([sys.executable, "-m", "tsm_benchmark_amd.codegen", *args],
                          capture_output=True, text=True)


def test_width_class():
    assert [codegen.width_class(x) for x in (1, 2, 3, 5, 8, 9, 33, 64)] == [1, 2, 4, 8, 8, 16, 64, 64]
    with pytest.raises(ValueError):
        codegen.width_class(65)
    with pytest.raises(ValueError):
        codegen.width_class(0)


def test_config_rules():
    ok = KernelConfig("f64", 8, 8, 8, 8, 256, 2)
    assert ok.problems() == [] and ok.ntiles == 1 and ok.row_lanes == 256
    assert KernelConfig("f64", 8, 8, 3, 8, 256, 2).problems()          # tile must divide
    assert KernelConfig("f64", 64, 64, 4, 4, 256, 2).problems()        # 256 tiles > 64
    assert KernelConfig("f64", 8, 8, 8, 8, 100, 2).problems()          # rows vs row lanes
    assert KernelConfig("f64", 8, 8, 8, 8, 256, 5).problems()          # pf
    assert KernelConfig("f64", 8, 8, 4, 4, 128, 4).problems() == []    # four in flight
    assert KernelConfig("f64", 64, 64, 8, 8, 256, 2).problems()        # LDS budget
    assert KernelConfig("bf16", 1, 1, 1, 1, 2, 1).problems()           # 16-byte chunks


def test_every_class_has_a_valid_default():
    for dt in codegen.DTYPES:
        for kc in codegen.CLASSES:
            for mc in codegen.CLASSES:
                c = codegen.default_config(dt, kc, mc)
                assert c.problems() == []
                assert c.lds_bytes <= codegen.LDS_BUDGET


def test_library_kernel_set():
    cfgs, defaults = codegen.library_kernels({"entries": []})
    names = [c.name for c in cfgs]
    assert len(names) == len(set(names))
    assert len(defaults) == 3 * 7 * 7
    # the reference's default set (M in {4, 8, 16}, TM in {2, 4}) is present for both depths
    for M in (4, 8, 16):
        for TM in (2, 4):
            for pf in (1, 2):
                assert any(c.dtype == "f64" and c.kc == M and c.tm == TM and c.tn == TM
                           and c.pf == pf for c in cfgs)


def test_tuning_entry_overrides_default():
    e = {"dtype": "f64", "kc": 8, "mc": 8, "tm": 4, "tn": 4, "rows": 128, "pf": 1, "wgs_per_cu": 3}
    cfgs, defaults = codegen.library_kernels({"entries": [e]})
    c, wpc, _ = defaults[("f64", 8, 8)]
    assert (c.tm, c.tn, c.rows, c.pf, wpc) == (4, 4, 128, 1, 3)
    with pytest.raises(ValueError):
        codegen.library_kernels({"entries": [dict(e, tm=3)]})


def test_emitted_kernel_text():
    c = KernelConfig("bf16", 32, 32, 8, 4, 256, 2)
    text = codegen.emit_kernel(c)
    assert "tsm_atb_bf16_32x32_t8x4_r256_p2(tsm::AtbArgs a)" in text
    assert "tsm::atb_body<tsm::bf16, 32, 32, 8, 4, 256, 2, 0>(a);" in text
    assert "hipLaunchKernelGGL(tsm_atb_bf16_32x32_t8x4_r256_p2" in text


def test_cli_rejects_inconsistent_flags():
    assert run_cli("--local-size", "128").returncode == 1
    assert run_cli("--sg-size", "32").returncode == 1
    assert run_cli("--unroll", "0").returncode == 1
    r = run_cli("--M", "8", "--tile-sizes", "4", "--stride", "1000")
    assert r.returncode == 1 and "whole number of work-groups" in r.stderr


def test_cli_name_layer():
    r = run_cli("--M", "4", "8", "--tile-sizes", "2", "4", "--name-layer", "-", "--with-benchmark")
    assert r.returncode == 0, r.stderr
    out = r.stdout
    for name in ("tsmttsm_4_4_2_2", "tsmttsm_4_4_4_4_lf", "tsmttsm_8_8_2_2_sg",
                 "tsmttsm_8_8_4_4_sg_lf"):
        assert f"inline tsm_status {name}(" in out
    assert "tsmttsm_4_4_8_8" not in out  # TM must divide M (sycl_codegen_new.py:645-647)
    assert "kGeneratedNames" in out
    # default stride follows the reference's rule with the gfx950 wavefront
    assert f"stride={codegen.suggest_stride([4, 8], [2, 4])}" in out


def test_tile_table_format_and_device_are_checked(tmp_path):
    """The tile table records its format, device and ROCm version; a table of
    another format or measured on another architecture is refused."""
    import json
    t = codegen.load_tuning()
    assert t["format"] == codegen.TUNING_FORMAT and t["device"].startswith("gfx950")
    assert "rocm" in t and "library_version" in t
    bad = dict(t, format=codegen.TUNING_FORMAT + 1)
    p = tmp_path / "t.json"
    p.write_text(json.dumps(bad))
    with pytest.raises(ValueError, match="format"):
        codegen.load_tuning(str(p))
    p.write_text(json.dumps(dict(t, device="gfx942:sramecc+:xnack-")))
    with pytest.raises(ValueError, match="gfx942"):
        codegen.load_tuning(str(p))
    # every tuned entry names a compiled kernel of the library
    cfgs, defaults = codegen.library_kernels(t)
    syms, sym_defaults = codegen.sym_kernels(defaults, t)
    for e in t["entries"]:
        c = defaults[(e["dtype"], e["kc"], e["mc"])][0]
        assert (c.tm, c.tn, c.rows, c.pf, c.nt) == (e["tm"], e["tn"], e["rows"], e["pf"], e["nt"])
    for e in t.get("sym_entries", []):
        c = sym_defaults[(e["dtype"], e["kc"])][0]
        assert c.sym == 1 and (c.rows, c.pf) == (e["rows"], e["pf"]) and c in syms


def test_direct_engine_rules_and_names():
    ok = KernelConfig("f64", 64, 64, 0, 0, 4, 8, 1, 0, 1)
    assert ok.problems() == [] and ok.engine == 3 and ok.mfma
    assert ok.name == "tsm_atb_f64_64x64_direct_d8_nt"
    assert ok.lds_bytes == 2 * 64 * 64 * 8
    assert KernelConfig("bf16", 64, 64, 0, 0, 4, 8, 1, 0, 1).problems()   # BF16 steps are 32 rows
    bf = KernelConfig("bf16", 64, 64, 0, 0, 32, 4, 1, 0, 1)
    assert bf.problems() == [] and bf.lds_bytes == 4 * 2 * 32 * 128 * 2
    assert KernelConfig("bf16", 64, 64, 0, 0, 32, 8, 1, 0, 1).problems()  # depth divides the fold
    assert KernelConfig("f64", 8, 8, 0, 0, 4, 8, 1, 0, 1).problems()      # multiples of 16
    assert KernelConfig("f32", 32, 32, 0, 0, 8, 8, 1, 0, 1).problems()    # steps are 4 rows
    assert KernelConfig("f32", 32, 32, 0, 0, 4, 6, 1, 0, 1).problems()    # depth divides 16
    assert KernelConfig("f64", 32, 32, 8, 8, 4, 8, 1, 0, 1).problems()    # no register tile
    assert KernelConfig("f64", 32, 32, 0, 0, 4, 8, 1, 1, 1).name.startswith("tsm_ata_f64_32x32_direct")
    assert [c.pf for c in codegen.direct_variants("f32", 16, 64)] == list(codegen.DIRECT_DEPTHS)
    assert [c.pf for c in codegen.direct_variants("bf16", 32, 32)] == list(codegen.DIRECT_DEPTHS_BF16)
    assert codegen.direct_variants("f64", 8, 16) == []


def test_direct_defaults_and_fallbacks():
    """Untuned FP64/FP32 classes of 1024+ outputs default to the direct engine
    with the staged heuristic kernel as fallback; a recorded fallback wins."""
    cfgs, defaults = codegen.library_kernels({"entries": []})
    fbs = codegen.library_fallbacks({"entries": []}, defaults)
    for dt in ("f64", "f32"):
        for kc, mc in ((32, 32), (16, 64), (64, 64)):
            c, wpc, il = defaults[(dt, kc, mc)]
            assert c.direct and wpc == 1 and il == 0 and c in cfgs
            fb, fil = fbs[(dt, kc, mc)]
            assert not fb.direct and fb == codegen.default_config(dt, kc, mc) and fb in cfgs
    for key in (("bf16", 64, 64), ("f64", 16, 16), ("f32", 8, 8)):  # BF16: tuned entries only
        assert not defaults[key][0].direct and fbs[key][0] == defaults[key][0]
    fb = {"tm": 0, "tn": 0, "rows": 64, "pf": 2, "nt": 1, "interleave": 1}
    e = {"dtype": "f64", "kc": 64, "mc": 64, "tm": 0, "tn": 0, "rows": 4, "pf": 16, "nt": 1,
         "direct": 1, "fallback": fb}
    t = {"entries": [e]}
    cfgs, defaults = codegen.library_kernels(t)
    c = defaults[("f64", 64, 64)][0]
    assert c.direct and c.pf == 16
    f, fil = codegen.library_fallbacks(t, defaults)[("f64", 64, 64)]
    assert (f.rows, f.pf, f.nt, f.direct, fil) == (64, 2, 1, 0, 1) and f in cfgs
    # A^T A: the direct default gives a direct A^T A kernel and its staged fallback
    syms, sym_defaults = codegen.sym_kernels(defaults, t)
    s = sym_defaults[("f64", 64)][0]
    assert s.direct and s.sym and codegen.sym_of(f) in syms
    # the generated table names the fallback beside the default
    table = codegen.emit_table(cfgs, defaults, t)
    assert "tsm_atb_f64_64x64_direct_d16_nt (else tsm_atb_f64_64x64_mfma_r64_p2_nt)" in table
    assert f'launch_{c.name}, 3}}' in table


def test_batched_classes_are_staged_with_wide_tiles():
    cfgs, defaults = codegen.library_kernels(codegen.load_tuning())
    for dt in codegen.DTYPES:
        for c in codegen.batched_classes(defaults, dt, codegen.load_tuning()):
            assert not c.direct and c.problems() == []
            assert c.lds_bytes <= codegen.BATCH_LDS
            if dt != "bf16":
                assert c.stage_bytes <= codegen.BATCH_STAGE
