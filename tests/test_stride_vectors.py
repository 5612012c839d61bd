"""The generator's stride rule and stride check against vectors written by the
reference's own generator (tests/golden/make_stride_vectors.py over
sycl_codegen_new.py:530-580), for codegen.py and for the oracle's copy."""
import json
import os

import oracle
from tsm_benchmark_amd import codegen

VEC = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "stride_vectors.json")))


def test_vectors_cover_both_outcomes():
    oks = [row[-1] for row in VEC["validate_stride"]]
    assert 0 < sum(oks) < len(oks)
    assert len(VEC["suggest_stride"]) >= 1000


def test_suggest_stride_matches_the_reference():
    for Ms, tiles, ls, sg, tw, want in VEC["suggest_stride"]:
        assert codegen.suggest_stride(Ms, tiles, local_size=ls, sg_size=sg, target_wgs=tw) == want, \
            (Ms, tiles, ls, sg, tw)
        assert oracle.suggest_stride(Ms, tiles, ls, sg, tw) == want, (Ms, tiles, ls, sg, tw)


def test_validate_stride_matches_the_reference():
    for Ms, tiles, ls, sg, stride, ok in VEC["validate_stride"]:
        msgs = codegen.validate_stride(Ms, tiles, stride, local_size=ls, sg_size=sg)
        assert (not msgs) == bool(ok), (Ms, tiles, ls, sg, stride, msgs)
        assert oracle.validate_stride(Ms, tiles, stride, ls, sg) == bool(ok), (Ms, tiles, ls, sg, stride)
