"""Writes tests/golden/ref_scores.json: scores of the reference libvship
(oracle/_ref/libvship_ref.so, built from the reference's own sources by
oracle/build_ref.sh) on seeded synthetic inputs that tests regenerate
bit-identically (vship_amd.synth / tests.cases). Runs on a GPU machine:

    python tests/golden/make_golden.py [out.json]
"""
from __future__ import annotations

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests import cases, refvship  # noqa: E402
from tests.golden.inputs import golden_inputs  # noqa: E402


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(ROOT, "tests", "golden", "ref_scores.json")
    ref = refvship.RefVship()
    records = []
    for name, cs1, p1, s1, cs2, p2, s2 in golden_inputs():
        score = ref.score(cs1, cs2, p1, s1, p2, s2)
        again = ref.score(cs1, cs2, p1, s1, p2, s2)  # the reference runs scales on two streams
        print(f"{name}: {score!r} (repeat {'equal' if again == score else repr(again)})", flush=True)
        records.append({"name": name, "reference_score": score, "repeat_equal": again == score})
    with open(out, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "library": "reference libvship 4.1.0 (gfx950 build)",
                   "cases": records}, f, indent=1)
    print("wrote", out, flush=True)
    _ = cases


if __name__ == "__main__":
    main()
