"""Writes tests/golden/stride_vectors.json: input/output vectors of the
reference generator's stride rule and stride check
(sycl_codegen_new.py:530-580, `suggest_stride` and `validate_stride`) over a
grid of --M, --tile-sizes, --local-size, --sg-size and target work-group
counts.  Run in a container that holds the reference checkout:

    python tests/golden/make_stride_vectors.py /root/reference

It imports the reference's generator module (standard library only) and
records its return values; nothing of the module is kept.  The tests compare
tsm_benchmark_amd.codegen and the oracle against these vectors."""
import contextlib
import importlib.util
import io
import itertools
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def load(ref_dir):
    spec = importlib.util.spec_from_file_location("ref_codegen",
                                                  os.path.join(ref_dir, "sycl_codegen_new.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main(ref_dir):
    ref = load(ref_dir)
    m_lists = [[4, 8, 16], [8], [1], [2, 4], [3, 6, 12], [4, 8, 16, 32], [5, 10], [64], [16, 32, 64],
               [6, 9]]
    tile_lists = [[2, 4], [1], [2], [4], [1, 2, 4, 8], [3], [8]]
    locals_ = [64, 128, 256, 512]
    sgs = [32, 64]
    targets = [2048, 100, 1]
    suggest, validate = [], []
    for Ms, tiles, ls, sg in itertools.product(m_lists, tile_lists, locals_, sgs):
        for tw in targets:
            s = ref.suggest_stride(Ms, tiles, ls, sg, tw)
            suggest.append([Ms, tiles, ls, sg, tw, s])
        base = ref.suggest_stride(Ms, tiles, ls, sg, 2048)
        for stride in sorted({base, base + ls, 524288, 3 * ls, ls * sg, 1000}):
            with contextlib.redirect_stderr(io.StringIO()):
                ok = ref.validate_stride(Ms, tiles, stride, ls, sg)
            validate.append([Ms, tiles, ls, sg, stride, int(bool(ok))])
    out = {"source": "sycl_codegen_new.py:530-580 (suggest_stride, validate_stride)",
           "suggest_stride_columns": ["M", "tiles", "local_size", "sg_size", "target_wgs", "stride"],
           "suggest_stride": suggest,
           "validate_stride_columns": ["M", "tiles", "local_size", "sg_size", "stride", "ok"],
           "validate_stride": validate}
    with open(os.path.join(HERE, "stride_vectors.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")
    print(f"{len(suggest)} suggest_stride and {len(validate)} validate_stride vectors")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
