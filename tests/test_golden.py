"""The FP64 oracle against scores of the reference libvship itself
(tests/golden/ref_scores.json, written on an MI355X by
tests/golden/make_golden.py from the reference built out of its own sources).

The inputs are regenerated here bit-identically from their seeds. Bars:
natural-content frames (sRGB 256x256, YUV 4:2:0 360p to 1080p, 8 and 10 bit)
within 0.005 (measured: at most 0.0026); 131x75 uniform-random planes within
0.1 (measured up to 0.083: the reference's float E[x^2] - mu^2 loses digits
at the 9x5 and 17x10 pyramid levels of such images, while the GPU path of
this library, which centres the moments, stays within 0.005 of the oracle on
the same inputs); the identical pair is exactly 100 in the oracle while the reference
scores it 99.91 (its numerator and denominator round differently for a == b).
"""
import json
import os

import pytest

import oracle
from tests.golden.inputs import golden_inputs

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "ref_scores.json")))
REF = {c["name"]: c["reference_score"] for c in GOLDEN["cases"]}


def _cases():
    # 1080p oracle runs take ~1.4 s each; keep the CPU suite to a few seconds
    for item in golden_inputs():
        name = item[0]
        if "1920x1080" in name and not name.endswith("frame0"):
            continue
        yield item


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_oracle_matches_reference_scores(case):
    name, cs1, p1, s1, cs2, p2, s2 = case
    got, _ = oracle.ssimu2(cs1, p1, s1, cs2, p2, s2)
    want = REF[name]
    if name.endswith("identity"):
        assert got == 100.0 and abs(want - 100.0) < 0.1
    elif name.startswith("random"):
        assert abs(got - want) <= 0.1, (got, want)
    else:
        assert abs(got - want) <= 0.005, (got, want)
