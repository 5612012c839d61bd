"""bench.py's contract on the CPU: one GPU only, and the CPU-baseline record
(the reference's kernels restated for host cores) on a small sample."""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)


def test_more_than_one_gpu_is_refused():
    for args, env in ((["--gpus", "2"], {}), ([], {"WORLD_SIZE": "4"})):
        r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args],
                           capture_output=True, text=True, timeout=120,
                           env=dict(os.environ, **env))
        assert r.returncode == 2 and "one GPU" in r.stderr
        assert r.stdout == ""


def test_cpu_baseline_record():
    import bench

    rec = bench.cpu_baseline(1 << 14, 8, 8, budget_s=2.0)
    assert rec["kind"] == "port" and rec["unit"] == "GB/s" and rec["cores"] >= 1
    assert rec["value"] > 0 and rec["value"] == max(rec["variants_GBps"].values())
    names = list(rec["variants_GBps"])
    assert len(names) == 3 + 3 * 3 * 2  # generated basic/_lf/_sg and tsmttsm4/6/9 x 3 strides x 2 tiles
    assert any(n.startswith("tsmttsm9") for n in names)
    json.dumps(rec)
