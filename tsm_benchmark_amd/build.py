"""Build libtsm.so (all HIP kernels for gfx950 plus the C ABI) in-tree.

Steps: run the generator (codegen.py --library), compile every translation
unit with hipcc in parallel into ``_build/``, link ``libtsm.so`` next to this
file.  Also used by ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

from . import codegen

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtsm.so")
BENCH_BIN = os.path.join(HERE, "tsm_bench")
CLOCK_BIN = os.path.join(HERE, "tsm_clock")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: keep hipcc from forming v_pk_{add,mul,fma}_f32 (the host
# pass warns that it does not know the feature; that warning is harmless).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"), "-I", CSRC,
]
HAND_WRITTEN = ["tsm_capi.hip", "tsm_backward.hip", "tsm_batched_host.hip"]


def sources() -> list[str]:
    gen = codegen.generate_library()
    return [os.path.join(CSRC, s) for s in HAND_WRITTEN] + gen


# Direct-engine kernels keep MFMA accumulators in ordinary VGPRs: in the
# default AGPR form the compiler copies them between the register files on
# every iteration of the streaming loop.
DIRECT_FLAGS = ["-mllvm", "-amdgpu-mfma-vgpr-form=true"]


def _compile(src: str) -> tuple[str, str]:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    extra = DIRECT_FLAGS if os.path.basename(src).startswith("tsm_direct_") else []
    cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")
    return obj, r.stderr


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compile every kernel and link libtsm.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    # largest translation units first so that the pool stays busy to the end
    srcs = sorted(sources(), key=lambda p: -os.path.getsize(p))
    jobs = jobs or min(16, os.cpu_count() or 4)
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = [o for o, _ in ex.map(_compile, srcs)]
    tmp = LIB + ".tmp"
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    # the C++ benchmark harness, linked against libtsm.so next to it
    harness = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(REPO, "include"),
               os.path.join(CSRC, "harness", "tsm_bench.cpp"), "-o", BENCH_BIN + ".tmp",
               "-L", HERE, "-ltsm", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(harness, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness build failed:\n{r.stderr[-4000:]}")
    os.replace(BENCH_BIN + ".tmp", BENCH_BIN)
    # the shader-clock timeline tool (a sampler kernel beside tsm_atb launches)
    clock = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-I", os.path.join(REPO, "include"),
             os.path.join(CSRC, "harness", "tsm_clock.hip"), "-o", CLOCK_BIN + ".tmp",
             "-L", HERE, "-ltsm", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(clock, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"tsm_clock build failed:\n{r.stderr[-4000:]}")
    os.replace(CLOCK_BIN + ".tmp", CLOCK_BIN)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources in {time.time() - t0:.1f}s", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
