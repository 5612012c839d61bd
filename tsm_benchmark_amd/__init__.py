"""tsm_benchmark_amd: tall-skinny C = A^T B on AMD Instinct MI355X (gfx950).

Layers:
  * ``libtsm.so`` - generated HIP kernels + the C ABI of include/tsm/tsm.h
    (built by ``python -m tsm_benchmark_amd.build``);
  * ``_lib``      - ctypes binding of that ABI (no torch);
  * ``ops``       - torch front end (``atb``, ``atb_autograd``, ``torch.ops.tsm.atb``);
  * ``codegen``   - the kernel generator; ``autotune`` - the tile-table tuner.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
