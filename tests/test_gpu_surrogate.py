"""GPU tests of the score surrogate (VshipAMD_MLP*): the HIP forward/backward
against the FP64 CPU network (oracle/mlp.py), the Adam update against its
closed form on the first step, and training on pairs labelled by the metric
itself, compared with the two predictors that use the same inputs without
the network (an affine least-squares fit, and the SSIMULACRA2 formula with
the scales the network does not see missing)."""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp
from vship_amd import ssimu2 as S
from vship_amd import surrogate as SG

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _case(n, seed=0):
    rng = np.random.default_rng(seed)
    p = SG.init_params(seed)
    # non-trivial biases and a non-degenerate last layer
    p[omlp.OFF_B1:omlp.OFF_W2] = rng.normal(0, 0.3, omlp.HID)
    p[omlp.OFF_B2:omlp.OFF_W3] = rng.normal(0, 0.3, omlp.HID)
    p[omlp.OFF_W3:omlp.OFF_B3] = rng.normal(0, 0.2, 2 * omlp.HID)
    p[omlp.OFF_B3:] = [0.1, -0.3]
    m = max(n, 64)
    X = rng.normal(0, 1, (m, omlp.IN)).astype(np.float32) * rng.uniform(0.01, 3, omlp.IN).astype(np.float32)
    X[:, 5] = 0.25  # constant column: ignored by the normalisation
    Y = rng.normal(65, 12, m).astype(np.float32)
    # normalisation from 64 samples, so a single sample does not normalise to 0
    xm = X[:64].mean(0).astype(np.float32)
    xs = X[:64].std(0).astype(np.float32)
    xs[5] = 0.0
    X, Y = X[:n], Y[:n]
    return p.astype(np.float32), X, Y, xm, xs, np.float32(65.0), np.float32(12.0)


@pytest.mark.parametrize("n", [1, 32, 77, 1000])
def test_gradients_match_cpu_fp64(n):
    p, X, Y, xm, xs, ym, ys = _case(n, seed=n)
    with SG.SurrogateNet(max_batch=1024) as net:
        net.set_params(p)
        net.set_norm(xm, xs, ym, ys)
        loss, g = net.gradients(X, Y)
    want_loss, want = omlp.loss_and_grad(p, X, Y, xm, xs, ym, ys)
    scale = np.abs(want).max()
    err = np.abs(g - want).max()
    _say(f"n={n}: loss gpu {loss:.7f} cpu {want_loss:.7f}; max |grad diff| {err:.2e} of max |grad| {scale:.2e}")
    assert loss == pytest.approx(want_loss, rel=2e-5)
    np.testing.assert_allclose(g, want, rtol=0, atol=2e-5 * scale + 1e-9)
    # every block of the gradient is populated (no parameter left out)
    for lo, hi in [(omlp.OFF_W1, omlp.OFF_B1), (omlp.OFF_B1, omlp.OFF_W2), (omlp.OFF_W2, omlp.OFF_B2),
                   (omlp.OFF_B2, omlp.OFF_W3), (omlp.OFF_W3, omlp.OFF_B3), (omlp.OFF_B3, omlp.PARAMS)]:
        assert np.abs(g[lo:hi]).max() > 0


def test_predict_matches_cpu_fp64():
    p, X, Y, xm, xs, ym, ys = _case(300, seed=5)
    with SG.SurrogateNet(max_batch=64) as net:
        net.set_params(p)
        net.set_norm(xm, xs, ym, ys)
        mean, sigma = net.predict(X)
        assert np.array_equal(net.params(), p)
    wm, wsig = omlp.predict(p, X, xm, xs, ym, ys)
    np.testing.assert_allclose(mean, wm, rtol=0, atol=1e-4 * ys)
    np.testing.assert_allclose(sigma, wsig, rtol=1e-4)


def test_first_training_step_is_adam():
    """One step over n samples from zero Adam state: the update is
    -lr_eff * g / (|g| + eps) with lr_eff = lr * 0.05 (the cosine schedule
    at its end for a one-step run), g the mean-loss gradient."""
    p, X, Y, xm, xs, ym, ys = _case(200, seed=9)
    lr = 1e-3
    with SG.SurrogateNet(max_batch=256) as net:
        net.set_params(p)
        net.set_norm(xm, xs, ym, ys)
        _, g = net.gradients(X, Y)
        losses = net.train(X, Y, epochs=1, batch=256, lr=lr, seed=1)
        after = net.params()
    step = -(lr * 0.05) * g / (np.abs(g) + 1e-8)
    big = np.abs(g) > 1e-5
    np.testing.assert_allclose((after - p)[big], step[big], rtol=2e-3, atol=2.5e-7)  # one float32 ulp at 1
    assert losses[0] == pytest.approx(omlp.loss_and_grad(p, X, Y, xm, xs, ym, ys)[0], rel=1e-4)


def _truncated_formula(features):
    """The SSIMULACRA2 weighted sum and remap with the features of the scales
    outside the network's inputs missing (set to 0)."""
    mask = np.array([(SG.FEATURE_SCALES >> ((i % 36) // 6)) & 1 == 1 for i in range(108)])
    return np.array([oracle.final_score(np.where(mask, f, 0.0)) for f in features])


def test_training_on_metric_labels_beats_methods_without_the_network():
    sizes = ((128, 128), (192, 108), (256, 256), (320, 180))
    n, hold = 5000, 800
    # small frames so the pairs synthesise in seconds
    gen = SG.TrainingPairs(seed=21, sources_per_size=16, sizes=sizes, size_prob=(0.3, 0.3, 0.2, 0.2))
    feats, scores, pixels, _ = SG.generate(n, chunk=1024, workers=8, pairs=gen)
    X = SG.inputs_from_features(feats, pixels)
    y = scores.astype(np.float32)
    tr = n - hold
    with SG.SurrogateNet(max_batch=512) as net:
        net.fit_norm(X[:tr], y[:tr])
        net.set_params(SG.init_params(3))
        losses = net.train(X[:tr], y[:tr], epochs=120, batch=256, lr=3e-3, seed=3)
        rep = SG.evaluate(net, X[tr:], y[tr:], fallback_sigma=1.5)
        mean, sigma = net.predict(X[tr:])
    lin = SG.linear_baseline(X[:tr], y[:tr], X[tr:], y[tr:])
    trunc = np.abs(_truncated_formula(feats[tr:]) - y[tr:])
    _say(f"loss {losses[0]:.3f} -> {losses[-1]:.3f}; scores {y.min():.1f}..{y.max():.1f}")
    _say(f"holdout MAE: network {rep['mae']:.3f} (p99 {rep['p99_abs_error']:.3f}), affine {lin['mae']:.3f}, "
         f"formula on the input scales only {trunc.mean():.3f}; fallback rate {rep['fallback_rate']:.3f}, "
         f"MAE after fallback {rep['mae_after_fallback']:.3f}")
    assert losses[-1] < losses[0]
    assert rep["mae"] < 0.7 * lin["mae"]
    assert rep["mae"] < 0.5 * trunc.mean()
    # the uncertainty head ranks errors: the uncertain half errs more
    err = np.abs(mean - y[tr:])
    hi = sigma > np.median(sigma)
    assert err[hi].mean() > err[~hi].mean()


def test_surrogate_errors():
    with SG.SurrogateNet(max_batch=8) as net:
        with pytest.raises(S.VshipError):
            net.set_params(np.zeros(10, np.float32))  # wrong parameter count
        p, X, Y, xm, xs, ym, ys = _case(20)
        net.set_params(p)
        with pytest.raises(S.VshipError):
            net.gradients(X, Y)  # more samples than max_batch
        with pytest.raises(S.VshipError):
            net.set_norm(xm, xs, 0.0, 0.0)  # target std must be positive
    lib = SG._bind(S.load_library())
    bad = S.Handler()
    bad.id = 9999
    assert lib.VshipAMD_MLPFree(bad) != 0
