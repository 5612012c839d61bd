"""The C++ harness (the reference's main.cc and generated-benchmark roles) on
an MI355X (-m gpu): every suite runs to completion and every kernel passes
the reference's acceptance rule against the naive host loop over all K rows
(main.cc:14-41).  The harness exits 1 when any kernel fails that check."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BENCH = os.path.join(HERE, "..", "tsm_benchmark_amd", "tsm_bench")


def run(*args):
    if not os.path.exists(BENCH):
        pytest.fail("tsm_bench is not built (python -m tsm_benchmark_amd.build)")
    r = subprocess.run([BENCH, *args, "--json"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, f"tsm_bench {' '.join(args)}: rc {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    rows = json.loads(r.stdout[r.stdout.index("["):])
    assert rows and all(row["correct"] for row in rows)
    return rows


def test_main_suite():
    """tsmttsm4 / 6 / 9 over main.cc's 18 configurations (M in {4, 8, 16})."""
    rows = run("--suite", "main", "--K", "65536", "--warmup", "1", "--runs", "2")
    assert len(rows) == 3 * 18


def test_generated_suite():
    """Every generated reference name tsmttsm_{M}_{M}_{TM}_{TM}[_lf|_sg|_sg_lf]."""
    rows = run("--suite", "generated", "--K", "65536", "--warmup", "1", "--runs", "2")
    names = {row["kernel"] for row in rows}
    assert "tsmttsm_8_8_4_4" in names and "tsmttsm_16_16_2_2_sg_lf" in names


@pytest.mark.parametrize("dtype,M,N", [("f64", 64, 64), ("f32", 64, 64), ("bf16", 64, 64),
                                       ("f64", 4, 32), ("f64", 8, 8)])
def test_atb_suite(dtype, M, N):
    """The library entry point (direct engine for the 64-wide FP64/FP32
    classes, staged kernels otherwise) on a ragged K."""
    rows = run("--suite", "atb", "--dtype", dtype, "--M", str(M), "--N", str(N), "--K", "100003",
               "--warmup", "1", "--runs", "3")
    assert len(rows) == 1
