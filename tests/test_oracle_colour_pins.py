"""Closed-form and published-standard pins of the oracle's colour branches
(CPU): HLG, ICtCp, BT.2020 constant luminance, ST 240, ST 428 and the
primaries matrices. Each check comes from a source other than the oracle's
own formula: the BT.2100 / BT.2020 / BT.709 / SMPTE publications (integer
matrices, primary chromaticities, white points, HLG constants) or an
algebraic property the reference's formula must have (luminance identity of
constant-luminance decoding, values at 0 and 1, odd symmetry).
"""
import numpy as np
import pytest

import oracle
from vship_amd import ssimu2 as S


def _lin(v, matrix, trc):
    return oracle.yuv_to_linear(*[float(x) for x in v], matrix, trc)


def _composite(matrix, trc):
    """3x3 matrix of the oracle's decode for a linear transfer (columns: unit inputs)."""
    return np.stack([_lin(e, matrix, trc) for e in np.eye(3)], 1)


# ---------------------------------------------------------------- HLG

HLG_A = 0.17883277
HLG_B = 1.0 - 4.0 * HLG_A
HLG_C = 0.5 - HLG_A * np.log(4.0 * HLG_A)  # BT.2100 Table 5


def _hlg_inverse_oetf(e):
    e = np.asarray(e, dtype=np.float64)
    return np.where(e <= 0.5, e * e / 3.0, (np.exp((e - HLG_C) / HLG_A) + HLG_B) / 12.0)


def test_hlg_constants_and_branch_continuity():
    # the published constants make the inverse OETF continuous at 1/2 and 1 at E' = 1
    assert _hlg_inverse_oetf(1.0) == pytest.approx(1.0, abs=5e-8)  # 8-digit constants
    assert (np.exp((0.5 - HLG_C) / HLG_A) + HLG_B) / 12.0 == pytest.approx(1.0 / 12.0, abs=1e-9)


@pytest.mark.parametrize("e", [0.0, 0.1, 0.3, 0.5, 0.6, 0.75, 0.9, 1.0])
def test_hlg_grey_is_inverse_oetf_to_the_system_gamma(e):
    # grey: the OOTF luminance (0.2627, 0.6780, 0.0593 sums to 1) is the
    # signal itself, so each channel is invOETF(E')^1.2 (BT.2100 system gamma)
    got = _lin([e, e, e], S.MATRIX_RGB, S.TRC_HLG)
    np.testing.assert_allclose(got, _hlg_inverse_oetf(e) ** 1.2, rtol=1e-7, atol=1e-12)


def test_hlg_ootf_keeps_channel_ratios():
    rng = np.random.default_rng(1)
    for _ in range(20):
        v = rng.uniform(0.05, 1.0, 3)
        got = _lin(v, S.MATRIX_RGB, S.TRC_HLG)
        scene = _hlg_inverse_oetf(v)
        ys = 0.2627 * scene[0] + 0.6780 * scene[1] + 0.0593 * scene[2]
        np.testing.assert_allclose(got, scene * ys ** 0.2, rtol=1e-12)


# ---------------------------------------------------------------- ICtCp

# BT.2100 Table 6 / 7 integer matrices (/4096)
RGB_TO_LMS = np.array([[1688, 2146, 262], [683, 2951, 462], [99, 309, 3688]]) / 4096.0
LMS_TO_ICTCP = np.array([[2048, 2048, 0], [6610, -13613, 7003], [17933, -17390, -543]]) / 4096.0


def test_ictcp_neutral_decodes_to_equal_lms_prime():
    # I alone (Ct = Cp = 0) must give L' = M' = S' = I (the published matrix:
    # I = (L' + M') / 2, Ct and Cp are differences). Decoding through PQ and
    # through a linear transfer then differs by the factor PQ(I) / I in every
    # channel, whatever the LMS -> RGB rows are; any other L'M'S' would not.
    np.testing.assert_allclose(np.linalg.inv(LMS_TO_ICTCP) @ [0.5, 0, 0], [0.5] * 3, rtol=1e-6)
    for i in (0.2, 0.5, 0.8):
        lin = _lin([i, 0.0, 0.0], S.MATRIX_BT2100_ICTCP, S.TRC_LINEAR)
        pq = _lin([i, 0.0, 0.0], S.MATRIX_BT2100_ICTCP, S.TRC_PQ)
        np.testing.assert_allclose(pq / lin, oracle.transfer(i, S.TRC_PQ) / i, rtol=1e-6)


def test_ictcp_neutral_gives_equal_lms():
    # I alone (Ct = Cp = 0) decodes to L' = M' = S' = I (published matrix:
    # I = (L' + M') / 2, Ct and Cp are differences); the first RGB row of the
    # LMS inverse sums to 1, so R = I for a linear transfer
    for i in (0.1, 0.4, 0.8):
        np.testing.assert_allclose(np.linalg.inv(LMS_TO_ICTCP) @ [i, 0, 0], [i, i, i], rtol=1e-6)
        assert _lin([i, 0.0, 0.0], S.MATRIX_BT2100_ICTCP, S.TRC_LINEAR)[0] == pytest.approx(i, rel=2e-7)


def test_ictcp_first_rgb_row_is_the_published_inverse():
    comp = _composite(S.MATRIX_BT2100_ICTCP, S.TRC_LINEAR)
    lms_prime = np.linalg.inv(LMS_TO_ICTCP)
    np.testing.assert_allclose(comp[0] @ np.linalg.inv(lms_prime), np.linalg.inv(RGB_TO_LMS)[0], rtol=2e-6,
                               atol=2e-7)


# ---------------------------------------------------------------- BT.2020 constant luminance

KR, KB = 0.2627, 0.0593
KG = 1.0 - KR - KB


@pytest.mark.parametrize("trc", [S.TRC_BT709, S.TRC_SRGB, S.TRC_PQ])
def test_bt2020_cl_linear_luminance_identity(trc):
    # constant-luminance decoding: Kr R + Kg G + Kb B equals the linear light
    # of Y' itself, for any Cb, Cr
    rng = np.random.default_rng(2)
    for _ in range(30):
        y, cb, cr = rng.uniform(0.05, 0.95), rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2)
        r, g, b = _lin([y, cb, cr], S.MATRIX_BT2020_CL, trc)
        assert KR * r + KG * g + KB * b == pytest.approx(oracle.transfer(y, trc), rel=1e-12, abs=1e-15)


def test_bt2020_cl_zero_chroma_is_grey():
    for y in (0.1, 0.5, 0.9):
        np.testing.assert_allclose(_lin([y, 0.0, 0.0], S.MATRIX_BT2020_CL, S.TRC_BT709), [y ** 2.4] * 3,
                                   rtol=1e-12)
        # Cr = 0 leaves R' = Y', Cb = 0 leaves B' = Y'
        assert _lin([y, 0.1, 0.0], S.MATRIX_BT2020_CL, S.TRC_BT709)[0] == pytest.approx(y ** 2.4, rel=1e-12)
        assert _lin([y, 0.0, 0.1], S.MATRIX_BT2020_CL, S.TRC_BT709)[2] == pytest.approx(y ** 2.4, rel=1e-12)


# ---------------------------------------------------------------- ST 240 and ST 428

def test_st240_closed_forms():
    f = lambda v: oracle.transfer(v, S.TRC_ST240_M)  # noqa: E731
    assert f(1.0) == pytest.approx(1.0, abs=1e-15)  # 1.1115 - 0.1115
    assert f(0.0) == 0.0
    assert f(0.01) == pytest.approx(0.04, rel=1e-15)  # linear segment 4 v below 0.0228
    assert f(0.5) == pytest.approx(1.1115 * 0.5 ** 0.45 - 0.1115, rel=1e-15)
    # the two segments meet at 0.0228 (within the 4-digit constants)
    assert abs(f(0.0228) - f(0.0227999)) < 2e-4
    for v in (0.01, 0.3, 0.9):
        assert f(-v) == -f(v)


def test_st428_closed_forms():
    f = lambda v: oracle.transfer(v, S.TRC_ST428)  # noqa: E731
    assert f(1.0) == pytest.approx(52.37 / 48.0, rel=1e-15)
    assert f(0.5) == pytest.approx(0.5 ** 2.6 * 52.37 / 48.0, rel=1e-14)
    assert f(0.0) == 0.0 and f(-0.5) == -f(0.5)


# ---------------------------------------------------------------- primaries

def _to_xyz(prim):
    """RGB -> XYZ of a primaries set, from the oracle's conversion to BT.709
    and its XYZ -> BT.709 matrix (the INTERNAL primaries)."""
    from_xyz = np.stack([oracle.primaries_to_709(e, S.PRIM_INTERNAL) for e in np.eye(3)], 1)
    to_709 = np.stack([oracle.primaries_to_709(e, prim) for e in np.eye(3)], 1)
    return np.linalg.inv(from_xyz) @ to_709


def _xy(col):
    return col[:2] / col.sum()


D65 = (0.3127, 0.3290)
ILLUMINANT_C = (0.3101, 0.3162)
# published primaries (R, G, B chromaticities) and white point
PUBLISHED = [
    (S.PRIM_BT709, ((0.64, 0.33), (0.30, 0.60), (0.15, 0.06)), D65, 1e-3),          # BT.709
    (S.PRIM_BT2020, ((0.708, 0.292), (0.170, 0.797), (0.131, 0.046)), D65, 2e-3),   # BT.2020
    (S.PRIM_BT470_M, ((0.67, 0.33), (0.21, 0.71), (0.14, 0.08)), ILLUMINANT_C, 2e-3),  # BT.470 System M
    (S.PRIM_BT470_BG, ((0.64, 0.33), (0.29, 0.60), (0.15, 0.06)), D65, 2e-3),      # BT.470 System B, G
    (S.PRIM_ST170_M, ((0.630, 0.340), (0.310, 0.595), (0.155, 0.070)), D65, 2e-4),  # SMPTE 170M
    (S.PRIM_ST240_M, ((0.630, 0.340), (0.310, 0.595), (0.155, 0.070)), D65, 2e-4),  # SMPTE 240M
]


@pytest.mark.parametrize("prim,rgb,white,tol", PUBLISHED, ids=lambda p: str(p))
def test_primaries_matrices_reproduce_published_chromaticities(prim, rgb, white, tol):
    m = _to_xyz(prim)
    for k in range(3):
        np.testing.assert_allclose(_xy(m[:, k]), rgb[k], atol=tol)
    w = m @ np.ones(3)
    np.testing.assert_allclose(_xy(w), white, atol=tol)
    assert w[1] == pytest.approx(1.0, abs=2e-3)  # white has luminance 1


def test_xyz_to_709_is_the_inverse_of_709_to_xyz():
    from_xyz = np.stack([oracle.primaries_to_709(e, S.PRIM_INTERNAL) for e in np.eye(3)], 1)
    to_xyz = np.array([[0.4124, 0.3576, 0.1805], [0.2126, 0.7152, 0.0722], [0.0193, 0.1192, 0.9505]])
    np.testing.assert_allclose(from_xyz @ to_xyz, np.eye(3), atol=1e-3)
