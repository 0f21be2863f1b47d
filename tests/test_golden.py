"""The oracle against scores of the reference libvship itself
(tests/golden/ref_scores.json, written on an MI355X by
tests/golden/make_golden.py from the reference built out of its own sources;
the inputs are regenerated here bit-identically from their seeds).

Two oracles take part. oracle.ssimu2 is the exact formula in FP64.
oracle.reference_float (oracle/ref_float.c) emulates the reference's own FP32
arithmetic (its float blur chains, uncentred E[x^2] - mu^2 moments with the
compiler's contractions, float block trees and float weighted sum) on the
FP64 colour front-end rounded to float. Bars, per kind of input:

- natural content (sRGB 256x256, YUV 4:2:0 360p to 1080p, 8 and 10 bit,
  the 16 frames of the throughput comparison): the emulation lands within
  0.002 of libvship, so libvship's distance from the exact formula (up to
  0.0084, frame 11) is its own float error; the exact oracle is within 0.01;
- the identical pair: libvship scores 99.9116, the emulation the same
  within 0.002, the exact formula exactly 100;
- the colourspace cases (131x75 uniform-random and 320x180 sinusoid content
  in every supported colourspace): there the reference's formula is
  ill-conditioned: moving half of the XYB values by one float ulp (the size
  of a libm difference between the reference's device cbrtf/powf and any
  other) moves its score with a standard deviation sigma of 0.001 to 0.024,
  measured here per case. Any two implementations of the formula (libvship,
  the emulation, the exact oracle) are then only defined to within that
  noise, so the bar is 0.002 + 6 sigma per case.

test_reference_scores_select_each_branch_reading checks the readings of the
reference's unusual colour branches against its own scores: each reading
lands within the case bar and a plausible alternative misses by more than 5
times that bar (30 sigma).
"""
import concurrent.futures as cf
import json
import os

import numpy as np
import pytest

import oracle
from tests.golden.inputs import golden_inputs
from vship_amd import ssimu2 as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "ref_scores.json")))
REF = {c["name"]: c["reference_score"] for c in GOLDEN["cases"]}
CASES = {item[0]: item for item in golden_inputs()}


def _colourspace_case(name):
    return name.startswith(("random131x75_", "smooth320x180_"))


def _scores(item):
    name, cs1, p1, s1, cs2, p2, s2 = item
    exact, _ = oracle.ssimu2(cs1, p1, s1, cs2, p2, s2)
    emul, _ = oracle.reference_float(cs1, p1, s1, cs2, p2, s2)
    return name, exact, emul


def test_fixture_covers_every_generated_input():
    assert set(REF) == set(CASES)
    assert all(c["repeat_equal"] for c in GOLDEN["cases"])  # the reference is deterministic


def test_natural_content_reference_error_is_its_float_arithmetic():
    natural = [item for name, item in CASES.items() if not _colourspace_case(name) and not name.endswith("identity")]
    assert len(natural) == 25
    # C calls release the GIL: one 1080p pair per thread
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_scores, natural))
    worst_emul = worst_exact = 0.0
    for name, exact, emul in results:
        ref = REF[name]
        assert abs(emul - ref) <= 0.002, (name, emul, ref)
        assert abs(exact - ref) <= 0.01, (name, exact, ref)
        worst_emul, worst_exact = max(worst_emul, abs(emul - ref)), max(worst_exact, abs(exact - ref))
    # the reference's float error is real and the emulation explains it
    assert worst_exact > 0.005 and worst_emul <= 0.002


def test_identical_pair():
    name, exact, emul = _scores(CASES["srgb256_seed0_identity"])
    assert exact == 100.0
    assert REF[name] < 99.95  # the reference's float numerator and denominator differ for a == b
    assert abs(emul - REF[name]) <= 0.002


def _noise_sigma(item, draws=12):
    name, cs1, p1, s1, cs2, p2, s2 = item
    r1 = oracle.linear_rgb(cs1, p1, s1).astype(np.float32)
    r2 = oracle.linear_rgb(cs2, p2, s2).astype(np.float32)
    return float(oracle.reference_float_spread(r1, r2, draws=draws, frac=0.5).std())


@pytest.mark.parametrize("name", [n for n in CASES if _colourspace_case(n)])
def test_colourspace_case_within_the_reference_float_noise(name):
    item = CASES[name]
    _, exact, emul = _scores(item)
    sigma = _noise_sigma(item)
    bar = 0.002 + 6.0 * sigma
    ref = REF[name]
    assert sigma < 0.03, sigma
    assert abs(exact - ref) <= bar, (exact, ref, sigma)
    assert abs(emul - ref) <= bar, (emul, ref, sigma)


# ---------------------------------------------------------------- branch readings

M2020 = np.array([[1.6605, -0.5876, -0.0728], [-0.1246, 1.1329, -0.0083], [-0.0182, -0.1006, 1.1187]])


def _apply(m, rgb):
    return np.einsum("ij,jhw->ihw", m, rgb)


def _prim_matrix(prim):
    return np.stack([oracle.primaries_to_709(e, prim) for e in np.eye(3)], 1)


def _gamma_rgb(cs, planes, strides):
    """R'G'B' (the decoded, upsampled and matrixed values before the transfer)."""
    c = S.Colorspace.from_buffer_copy(cs)
    c.transferFunction = S.TRC_LINEAR
    c.primaries = S.PRIM_BT709
    return oracle.linear_rgb(c, planes, strides)


def _emulated(rgb1, rgb2):
    return oracle.reference_float_from_linear(rgb1.astype(np.float32), rgb2.astype(np.float32))[0]


def _pair(name):
    _, cs1, p1, s1, cs2, p2, s2 = CASES[name]
    return (cs1, p1, s1), (cs2, p2, s2)


def _ours_and_alternative(name, alternative):
    a, b = _pair(name)
    ours = _emulated(oracle.linear_rgb(*a), oracle.linear_rgb(*b))
    alt = _emulated(alternative(*a), alternative(*b))
    return ours, alt


def _st240_smpte_eotf(cs, planes, strides):
    # SMPTE 240M's decoding (inverse of its OETF), where the reference applies
    # the OETF's shape 1.1115 V^0.45 - 0.1115 itself (transferToLinear.hpp:143-158)
    v = _gamma_rgb(cs, planes, strides)
    av = np.abs(v)
    lin = np.where(av < 0.0913, av / 4.0, ((av + 0.1115) / 1.1115) ** (1.0 / 0.45)) * np.sign(v)
    return _apply(_prim_matrix(cs.primaries), lin)


def _st428_without_scale(cs, planes, strides):
    return oracle.linear_rgb(cs, planes, strides) * (48.0 / 52.37)


def _pq_normalised(cs, planes, strides):
    # PQ normalised to 1 at 10000 cd/m2 instead of the reference's Lmax = 100
    return oracle.linear_rgb(cs, planes, strides) / 100.0


def _hlg_without_ootf(cs, planes, strides):
    # BT.2020 CL + HLG: divide out the OOTF gain k = Ys^0.2 applied to (R, Y, B)
    rgb = _apply(np.linalg.inv(M2020), oracle.linear_rgb(cs, planes, strides))
    y = 0.2627 * rgb[0] + 0.6780 * rgb[1] + 0.0593 * rgb[2]
    ys_final = 0.2627 * rgb[0] + 0.6780 * y + 0.0593 * rgb[2]
    k = np.maximum(ys_final, 1e-12) ** (0.2 / 1.2)
    return _apply(M2020, rgb / k)


def _cl_as_ncl(cs, planes, strides):
    c = S.Colorspace.from_buffer_copy(cs)
    c.YUVMatrix = S.MATRIX_BT2020_NCL
    return oracle.linear_rgb(c, planes, strides)


def _ictcp_textbook_rows(cs, planes, strides):
    # BT.2100's LMS -> RGB inverse where the reference evaluates its rows in
    # place, each using the components already overwritten (YUVToLinRGB.hpp:28-43)
    r = np.array([3.4366066943330784267, -2.5064521186562698975, 0.069845424323191470954])
    g = -0.23026480071281397 * r + np.array([0, 1.4064527541936944, -0.2083538389058436])
    b = -0.02684381778741865 * r - 0.08378524945770065 * g + np.array([0, 0, 1.106879231557486])
    in_place = np.stack([r, g, b])
    textbook = np.linalg.inv(np.array([[1688, 2146, 262], [683, 2951, 462], [99, 309, 3688]]) / 4096.0)
    a = M2020 @ textbook @ np.linalg.inv(in_place) @ np.linalg.inv(M2020)
    return _apply(a, oracle.linear_rgb(cs, planes, strides))


BRANCHES = [
    ("smooth320x180_yuv420p14_st240", _st240_smpte_eotf),
    ("random131x75_yuv420p8_st240trc", _st240_smpte_eotf),
    ("smooth320x180_yuv444p8_st428_xyz", _st428_without_scale),
    ("smooth320x180_yuv444p16_2020ncl_pq", _pq_normalised),
    ("smooth320x180_yuv420p12_2020cl_hlg", _hlg_without_ootf),
    ("smooth320x180_yuv420p12_2020cl_hlg", _cl_as_ncl),
    ("random131x75_yuv420p10_ictcp_pq", _ictcp_textbook_rows),
    ("smooth320x180_yuv420p10_ictcp_pq", _ictcp_textbook_rows),
]


@pytest.mark.parametrize("name,alternative", BRANCHES, ids=[f"{n}-{f.__name__}" for n, f in BRANCHES])
def test_reference_scores_select_each_branch_reading(name, alternative):
    ours, alt = _ours_and_alternative(name, alternative)
    bar = 0.002 + 6.0 * _noise_sigma(CASES[name])
    ref = REF[name]
    assert abs(ours - ref) <= bar, (ours, ref, bar)
    assert abs(alt - ref) > 5.0 * bar, (alt, ref, bar)
