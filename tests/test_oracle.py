"""Pins of the FP64 oracle (CPU): every check comes from a source other than
the oracle's own formula - published constants and closed forms, a textbook
library routine (scipy), the inverse of a forward matrix, brute-force
equivalents and invariants chosen so that a dropped term, a wrong sign or an
off-by-one index fails."""
import numpy as np
import pytest
from scipy import ndimage

import oracle
from vship_amd import ssimu2 as S
from vship_amd import synth
from tests import cases


def test_taps_match_published_values():
    # float32 tap values of sigma 1.5, radius 8 (center last)
    want = [1.77086775e-07, 4.96402845e-06, 8.92201497e-05, 0.00102818618, 0.00759732397,
            0.0359939784, 0.109340049, 0.212965325, 0.265961528]
    g = oracle.gauss_taps()
    np.testing.assert_allclose(g[:9], want, rtol=1e-6)  # float32-rounded values
    np.testing.assert_allclose(g[8:], want[::-1], rtol=1e-6)
    # a sampled unit Gaussian sums to 1 up to ~1e-19 aliasing terms, so the
    # 17 taps sum to 1 minus the two truncated tails |k| >= 9
    tail = sum(np.exp(-k * k / 4.5) / np.sqrt(2 * np.pi * 2.25) for k in range(9, 60))
    assert abs(g.sum() - (1.0 - 2.0 * tail)) < 1e-15
    assert g[8:].sum() == pytest.approx(0.632980751, rel=1e-8)


@pytest.mark.parametrize("h,w", [(23, 17), (5, 40), (1, 9), (30, 1)])
def test_blur_is_normalised_zero_padded_correlation(h, w):
    rng = np.random.default_rng(h * 100 + w)
    img = rng.uniform(0, 1, (h, w))
    g = oracle.gauss_taps()
    k2 = np.outer(g, g)
    num = ndimage.correlate(img, k2, mode="constant", cval=0.0)
    den = ndimage.correlate(np.ones_like(img), k2, mode="constant", cval=0.0)
    np.testing.assert_allclose(oracle.blur(img), num / den, rtol=1e-12, atol=1e-14)


def test_blur_of_constant_is_constant_everywhere():
    img = np.full((19, 33), 0.37)
    np.testing.assert_allclose(oracle.blur(img), 0.37, rtol=1e-14)


@pytest.mark.parametrize("h,w", [(7, 9), (8, 8), (1, 5), (5, 1)])
def test_downsample_equals_edge_padded_box_mean(h, w):
    rng = np.random.default_rng(h + 10 * w)
    img = rng.uniform(0, 1, (h, w))
    H, W = h + (h & 1), w + (w & 1)
    pad = np.pad(img, ((0, H - h), (0, W - w)), mode="edge")
    want = pad.reshape(H // 2, 2, W // 2, 2).mean(axis=(1, 3))
    np.testing.assert_allclose(oracle.downsample(img), want, rtol=1e-15)


def test_xyb_grey_axis_closed_forms():
    v = np.array([0.0, 0.1, 0.5, 1.0, 3.0])
    x, y, b = oracle.xyb(v, v, v)
    np.testing.assert_allclose(x, 0.42, atol=1e-15)
    np.testing.assert_allclose(b, 0.55, atol=1e-15)
    np.testing.assert_allclose(y[[0, 2, 3]], [0.01, 0.6497483124, 0.8553085600], atol=2e-10)


def test_xyb_clamps_negative_opsin():
    x, y, b = oracle.xyb([-5.0], [-5.0], [-5.0])
    # every opsin channel clamps to 0 -> cbrt(0) - cbrt(bias) mixes to 0
    assert abs(x[0] - 0.42) < 1e-15 and abs(b[0] - 0.55) < 1e-15
    assert abs(y[0] - (0.01 - 0.1559542025327239)) < 1e-12


def test_transfer_closed_forms():
    assert oracle.transfer(0.5, S.TRC_SRGB) == pytest.approx(((0.5 + 0.055) / 1.055) ** 2.4, rel=1e-14)
    assert oracle.transfer(0.04, S.TRC_SRGB) == pytest.approx(0.04 / 12.92, rel=1e-14)
    assert oracle.transfer(-0.5, S.TRC_SRGB) == pytest.approx(-oracle.transfer(0.5, S.TRC_SRGB), rel=1e-14)
    assert oracle.transfer(0.3, S.TRC_BT709) == pytest.approx(0.3 ** 2.4, rel=1e-14)
    assert oracle.transfer(-0.3, S.TRC_BT709) == pytest.approx(-(0.3 ** 2.4), rel=1e-14)
    assert oracle.transfer(0.3, S.TRC_BT470_M) == pytest.approx(0.3 ** 2.2, rel=1e-14)
    assert oracle.transfer(0.3, S.TRC_BT601) == pytest.approx(0.3 ** 2.8, rel=1e-14)
    assert oracle.transfer(1.0, S.TRC_PQ) == pytest.approx(100.0, rel=1e-12)
    assert oracle.transfer(0.0, S.TRC_PQ) == 0.0
    assert oracle.transfer(0.7, S.TRC_LINEAR) == 0.7


def test_range_closed_forms():
    u8, u10 = S.SAMPLE_UINT8, S.SAMPLE_UINT10
    L, F = S.RANGE_LIMITED, S.RANGE_FULL
    assert oracle.range_to_full(16, u8, L, False) == 0.0
    assert oracle.range_to_full(235, u8, L, False) == pytest.approx(1.0, abs=1e-15)
    assert oracle.range_to_full(128, u8, L, True) == 0.0
    assert oracle.range_to_full(240, u8, L, True) == pytest.approx(0.5, abs=1e-15)
    assert oracle.range_to_full(64, u10, L, False) == 0.0
    assert oracle.range_to_full(940, u10, L, False) == pytest.approx(1.0, abs=1e-15)
    assert oracle.range_to_full(255, u8, F, False) == 1.0
    assert oracle.range_to_full(1023, u10, F, True) == pytest.approx(0.5, abs=1e-15)
    assert oracle.range_to_full(0.25, S.SAMPLE_FLOAT, F, False) == 0.25


@pytest.mark.parametrize("matrix,kr,kb", [(S.MATRIX_BT709, 0.2126, 0.0722), (S.MATRIX_BT470_BG, 0.299, 0.114),
                                          (S.MATRIX_BT2020_NCL, 0.2627, 0.0593)])
def test_ncl_matrix_is_inverse_of_forward_matrix(matrix, kr, kb):
    kg = 1 - kr - kb
    fwd = np.array([[kr, kg, kb],
                    [-kr / (2 * (1 - kb)), -kg / (2 * (1 - kb)), 0.5],
                    [0.5, -kg / (2 * (1 - kr)), -kb / (2 * (1 - kr))]])
    inv = np.linalg.inv(fwd)
    rng = np.random.default_rng(1)
    for _ in range(5):
        yuv = rng.uniform(-0.5, 1, 3)
        got = oracle.yuv_to_linear(yuv[0], yuv[1], yuv[2], matrix, S.TRC_LINEAR)
        np.testing.assert_allclose(got, inv @ yuv, rtol=1e-12, atol=1e-12)


def test_primaries_identity_and_roundtrip():
    rgb = np.array([0.2, 0.5, 0.7])
    np.testing.assert_array_equal(oracle.primaries_to_709(rgb, S.PRIM_BT709), rgb)
    # XYZ of BT.709 white (D65, Y = 1) maps back to equal RGB
    xyz_white = np.array([0.9505, 1.0, 1.089])
    np.testing.assert_allclose(oracle.primaries_to_709(xyz_white, S.PRIM_INTERNAL), [1, 1, 1], atol=3e-3)


def _upsample_1d(sig, sub, loc):
    n_full = len(sig) << sub
    small = np.asarray(sig, dtype=np.float64)[None, :]
    return oracle.upsample_plane(small, n_full, 1, sub, 0, loc)[0]


def test_catmull_rom_published_weights():
    # Catmull-Rom kernel weights at t: (-t(1-t)^2, 2-5t^2+3t^3, t(1+4t-3t^2), -t^2(1-t)) / 2
    def w(t):
        return np.array([-t * (1 - t) ** 2, 2 - 5 * t ** 2 + 3 * t ** 3, t * (1 + 4 * t - 3 * t ** 2),
                         -t ** 2 * (1 - t)]) / 2
    for t in (0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875):
        for j in range(4):
            pts = [0.0] * 4
            pts[j] = 1.0
            assert oracle.catmull_rom(*pts, t) == pytest.approx(w(t)[j], abs=1e-15)
    np.testing.assert_allclose(w(0.5), [-1 / 16, 9 / 16, 9 / 16, -1 / 16])
    np.testing.assert_allclose(w(0.25), [-9 / 128, 111 / 128, 29 / 128, -3 / 128])


def test_chroma_upsampling_siting():
    imp = np.zeros(12)
    imp[6] = 1.0
    left = _upsample_1d(imp, 1, S.CHROMA_LEFT)
    # cosited: even outputs copy the samples, odd ones are t = 0.5 midpoints
    assert left[12] == 1.0 and left[10] == 0.0 and left[14] == 0.0
    np.testing.assert_allclose(left[[9, 11, 13, 15]], [-1 / 16, 9 / 16, 9 / 16, -1 / 16], atol=1e-15)
    centre = _upsample_1d(imp, 1, S.CHROMA_CENTER)
    # centred: sample k sits at 2k + 0.5; outputs at 2k+1 (t = .25), 2k+2 (t = .75)
    np.testing.assert_allclose(centre[[13, 12]], [111 / 128, 111 / 128], atol=1e-15)
    np.testing.assert_allclose(centre[[14, 11]], [29 / 128, 29 / 128], atol=1e-15)
    ramp = np.arange(12, dtype=np.float64)
    for sub, loc, off in [(1, S.CHROMA_LEFT, 0.0), (1, S.CHROMA_CENTER, 0.5), (2, S.CHROMA_LEFT, 0.0),
                          (2, S.CHROMA_CENTER, 1.5)]:
        out = _upsample_1d(ramp, sub, loc)
        x = np.arange(len(out))
        interior = (x >= 3 << sub) & (x < (len(ramp) - 3) << sub)
        np.testing.assert_allclose(out[interior], ((x - off) / (1 << sub))[interior], atol=1e-13)


def test_pruned_map_set_matches_weight_table():
    # (plane, scale) -> maps computed; X = 0, Y = 1, B = 2; maps ssim, art, det
    active = {(0, 1): "s", (0, 2): "sa", (0, 3): "sa", (0, 4): "a",
              (1, 0): "sd", (1, 1): "sd", (1, 2): "s", (1, 3): "sa", (1, 4): "sa",
              (2, 3): "sa", (2, 4): "sad"}
    for p in range(3):
        for s in range(6):
            got = "".join(c for m, c in enumerate("sad") if oracle.map_active(p, s, m))
            assert got == active.get((p, s), ""), (p, s)
    w = oracle.weights()
    assert w.shape == (108,) and abs(w.sum() - 888.31) < 0.01
    assert w[48] == pytest.approx(225.20515300849274) and w[100] == pytest.approx(171.2667255897307)


def test_final_score_closed_forms():
    w = oracle.weights()
    i = 48
    for S_, want in [(0.0, 100.0), (1.0, 83.5696784), (10.0, 33.6420343)]:
        f = np.zeros(108)
        f[i] = S_ / w[i]
        assert oracle.final_score(f) == pytest.approx(want, abs=1e-6)


def test_ssim_map_of_shifted_image_is_shift_squared():
    # b = a + c: means shift by c everywhere (renormalised blur), variances and
    # covariance are equal, so ssim = 1 - (1 - c^2) = c^2 and the edge maps are 0
    rng = np.random.default_rng(4)
    a = rng.uniform(0, 1, (21, 26))
    c = 0.03
    s, ar, d = oracle.maps(a, a + c)
    np.testing.assert_allclose(s, c * c, rtol=1e-6)
    assert np.abs(ar).max() < 1e-12 and np.abs(d).max() < 1e-12


def test_ssim_map_of_scaled_image():
    rng = np.random.default_rng(5)
    a = rng.uniform(0.2, 1, (15, 18))
    k = 0.9
    mu = oracle.blur(a)
    var = oracle.blur(a * a) - mu * mu
    C2 = 0.0009
    want = 1 - (1 - ((1 - k) * mu) ** 2) * (2 * k * var + C2) / ((1 + k * k) * var + C2)
    s, _, _ = oracle.maps(a, k * a)
    np.testing.assert_allclose(s, np.maximum(want, 0), rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("name,kw", cases.COLOR_CASES[:6] + cases.COLOR_CASES[16:18],
                         ids=[c[0] for c in cases.COLOR_CASES[:6] + cases.COLOR_CASES[16:18]])
def test_identity_is_exactly_100(name, kw):
    cs = S.colorspace(37, 23, **kw)
    ref, _, st = cases.perturbed_pair(cs, seed=2)
    score, feats = oracle.ssimu2(cs, ref, st, cs, ref, st)
    assert score == 100.0 and not feats.any()


def test_score_decreases_with_noise():
    ref, _ = synth.srgb_pair(96, 96, seed=9)
    cs = S.srgb_rgb_full(96, 96)
    rng = np.random.default_rng(0)
    noise = rng.normal(0, 1, (3, 96, 96))
    scores = []
    for sigma in (1, 3, 9, 27):
        dist = [np.clip(np.round(r + sigma * n), 0, 255).astype(np.uint8) for r, n in zip(ref, noise)]
        scores.append(oracle.ssimu2(cs, ref, [96] * 3, cs, dist, [96] * 3)[0])
    assert all(a > b for a, b in zip(scores, scores[1:])), scores
    assert scores[0] < 100


def test_unpruned_features_superset():
    ref, dist = synth.srgb_pair(64, 64, seed=2)
    cs = S.srgb_rgb_full(64, 64)
    s1, f1 = oracle.ssimu2(cs, ref, [64] * 3, cs, dist, [64] * 3, prune=True)
    s0, f0 = oracle.ssimu2(cs, ref, [64] * 3, cs, dist, [64] * 3, prune=False)
    nz = f1 != 0
    np.testing.assert_array_equal(f1[nz], f0[nz])
    # the dropped weights sum to 0.0133, so the scores stay close
    assert abs(s1 - s0) < 0.05


# ---- the surrogate network's FP64 restatement (oracle/mlp.py)

def _mlp_case(seed=0, n=7):
    from oracle import mlp
    rng = np.random.default_rng(seed)
    p = (rng.normal(0, 0.3, mlp.PARAMS)).astype(np.float32)
    X = rng.normal(0, 2, (n, mlp.IN)).astype(np.float32)
    Y = rng.normal(60, 15, n).astype(np.float32)
    xmean = X.mean(0).astype(np.float32)
    xstd = X.std(0).astype(np.float32)
    xstd[3] = 0.0  # a constant column is ignored
    return p, X, Y, xmean, xstd, np.float32(60.0), np.float32(15.0)


def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp
    p, X, Y, xm, xs, ym, ys = _mlp_case()
    _, g = mlp.loss_and_grad(p, X, Y, xm, xs, ym, ys, quantise=False)
    rng = np.random.default_rng(1)
    # every block of the parameter vector: weights and biases of each layer
    picks = []
    for lo, hi in [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                   (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.PARAMS)]:
        picks += list(rng.integers(lo, hi, 6))
    pd = p.astype(np.float64)
    for k in picks:
        h = 1e-3
        up, dn = pd.copy(), pd.copy()
        up[k] += h
        dn[k] -= h
        # float64 parameters through the unquantised path
        lu, _ = mlp.loss_and_grad(up, X, Y, xm, xs, ym, ys, quantise=False)
        ld, _ = mlp.loss_and_grad(dn, X, Y, xm, xs, ym, ys, quantise=False)
        fd = (lu - ld) / (2 * h)
        assert fd == pytest.approx(g[k], rel=2e-3, abs=1e-6), k


def test_mlp_oracle_gelu_and_loss_closed_forms():
    from math import erf, exp, pi, sqrt
    from oracle import mlp
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    want = [0.5 * v * (1 + erf(v / sqrt(2))) for v in z]
    np.testing.assert_allclose(mlp.gelu(z), want, rtol=1e-15)
    # derivative of x * Phi(x) is Phi(x) + x * phi(x); at 0 it is 1/2
    assert mlp.gelu_grad(np.array([0.0]))[0] == 0.5
    dz = 1e-6
    for v in z:
        fd = (mlp.gelu(np.array([v + dz])) - mlp.gelu(np.array([v - dz])))[0] / (2 * dz)
        assert mlp.gelu_grad(np.array([v]))[0] == pytest.approx(fd, rel=1e-7)
    assert mlp.gelu_grad(np.array([1.0]))[0] == pytest.approx(0.5 * (1 + erf(1 / sqrt(2))) + exp(-0.5) / sqrt(2 * pi))
    # all-zero weights: mu = lv = b3 = 0, so the loss is the mean of y^2 / 2
    p = np.zeros(mlp.PARAMS, np.float32)
    X = np.ones((4, mlp.IN), np.float32)
    Y = np.array([1.0, -2.0, 0.5, 3.0], np.float32)
    loss, g = mlp.loss_and_grad(p, X, Y, np.zeros(mlp.IN, np.float32), np.ones(mlp.IN, np.float32), 0.0, 1.0)
    assert loss == pytest.approx(np.mean(Y.astype(np.float64) ** 2) / 2)
    # d loss / d b3 = (-mean(y), mean(0.5 * (1 - y^2)))
    assert g[mlp.OFF_B3] == pytest.approx(-np.mean(Y))
    assert g[mlp.OFF_B3 + 1] == pytest.approx(np.mean(0.5 * (1 - Y.astype(np.float64) ** 2)))


def test_bf16_rounding_ties_to_even():
    from oracle import mlp
    # 1 + 2^-8 is halfway between bf16 1 and 1 + 2^-7: ties go to the even 1.0
    vals = np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], np.float32)
    np.testing.assert_array_equal(mlp.bf16_round(vals), np.array([1.0, 1.0 + 2 ** -6, 1.0, -2.5], np.float32))


# ---- crop and resize (reference gpuColorToLinear/resize.hpp:26-65, vshipColor.hpp:87-103)

def _linear_cs(w, h, tw=-1, th=-1, crop=(0, 0, 0, 0)):
    cs = S.colorspace(w, h, sample=S.SAMPLE_FLOAT, rng=S.RANGE_FULL, subw=0, subh=0, family=S.FAMILY_RGB,
                      matrix=S.MATRIX_RGB, trc=S.TRC_LINEAR, prim=S.PRIM_BT709)
    cs.target_width, cs.target_height = tw, th
    cs.crop = S.CropRectangle(*crop)  # top, bottom, left, right
    return cs


def _linear_planes(img):
    planes = [np.ascontiguousarray(p, dtype=np.float32) for p in img]
    return planes, [p.strides[0] for p in planes]


def test_resize_constant_and_integer_downscale():
    rng = np.random.default_rng(3)
    const = [np.full((30, 44), v, np.float32) for v in (0.2, 0.5, 0.9)]
    out = oracle.linear_rgb(_linear_cs(44, 30, 71, 13), *_linear_planes(const))
    np.testing.assert_allclose(out, np.stack([np.full((13, 71), v) for v in (0.2, 0.5, 0.9)]), rtol=1e-6)
    # halving: pos = 2x exactly, t = 0, so the output is every other sample
    img = [rng.uniform(0, 1, (40, 60)).astype(np.float32) for _ in range(3)]
    out = oracle.linear_rgb(_linear_cs(60, 40, 30, 20), *_linear_planes(img))
    np.testing.assert_allclose(out, np.stack([p[::2, ::2] for p in img]).astype(np.float64), rtol=0, atol=0)


def test_resize_reproduces_a_linear_ramp_in_the_interior():
    # Catmull-Rom reproduces linear functions where all four taps are inside
    w, h, tw = 20, 6, 50
    ramp = np.tile(0.1 + 0.03 * np.arange(w, dtype=np.float32), (h, 1))
    out = oracle.linear_rgb(_linear_cs(w, h, tw, h), *_linear_planes([ramp] * 3))
    pos = np.arange(tw) * w / tw
    inner = (pos >= 1) & (np.floor(pos) + 2 <= w - 1)
    np.testing.assert_allclose(out[0][:, inner], np.tile(0.1 + 0.03 * pos[inner], (h, 1)), rtol=1e-6)


def test_crop_is_a_window_of_the_resized_image():
    rng = np.random.default_rng(4)
    img = [rng.uniform(0, 1, (24, 32)).astype(np.float32) for _ in range(3)]
    full = oracle.linear_rgb(_linear_cs(32, 24, 48, 36), *_linear_planes(img))
    crop = oracle.linear_rgb(_linear_cs(32, 24, 48, 36, crop=(3, 5, 7, 2)), *_linear_planes(img))
    assert crop.shape == (3, 36 - 3 - 5, 48 - 7 - 2)
    np.testing.assert_array_equal(crop, full[:, 3:36 - 5, 7:48 - 2])
    # crop without resize is a plain window
    c2 = oracle.linear_rgb(_linear_cs(32, 24, crop=(1, 2, 3, 4)), *_linear_planes(img))
    np.testing.assert_array_equal(c2, np.stack(img)[:, 1:22, 3:28].astype(np.float64))
