"""GPU test of target-quality screening (vship_amd.surrogate.Screening) at
reduced size: 6 sources at 1920x1080 (8-bit YUV 4:2:0) x 16 DCT-quantised
candidates, target score 70. The surrogate-screened selection must equal the
exact-only selection; every score it confirms must equal the exact batched
score of the same candidate bit for bit (the confirmation completes the
scales outside the network's inputs on the planes the feature pass stored);
and a fallback threshold forced to 0 sends every candidate off the shortlist
to the exact metric and still agrees. Both shipped networks: inputs from
scales 0-2 and from scales 2-4."""
import numpy as np
import pytest

from vship_amd import ssimu2 as S
from vship_amd import surrogate as SG

pytestmark = pytest.mark.gpu

TARGET = 70.0


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def screening():
    refs, dists, qs = SG.screening_set(6, 16, seed=9)
    scr = SG.Screening(refs, dists, qs, 1920, 1080)
    yield scr
    scr.close()


@pytest.mark.parametrize("params,scales", [(SG.PARAMS_FILE, SG.FEATURE_SCALES),
                                           (SG.COARSE_PARAMS_FILE, SG.COARSE_SCALES)], ids=["scales012", "scales234"])
def test_surrogate_screening_agrees_with_exact_only(screening, params, scales):
    ncand = screening.nsrc * screening.ncand
    with SG.SurrogateNet(max_batch=ncand) as net, S.SSIMU2Batch(ncand, S.STORAGE_FP16) as bt:
        net.load(params)
        assert net.feature_scales == scales
        exact_choice, exact_scores = screening.exact_only(bt, TARGET)
        choice, stats, mean, sigma = screening.with_surrogate(bt, net, TARGET)
        _say("exact", exact_choice, "surrogate", choice, "confirmed", stats["confirmed_pairs"],
             "mae", float(np.abs(mean - exact_scores).mean()))
        assert choice == exact_choice
        assert stats["confirmed_pairs"] < ncand
        for (i, k), v in stats["confirmed"].items():
            assert v == exact_scores[i, k], (i, k, v, exact_scores[i, k])
        forced, fstats, _, _ = screening.with_surrogate(bt, net, TARGET, fallback_sigma=0.0)
        _say("forced fallback", fstats["uncertain_pairs_sent_to_exact"], fstats["confirmed_pairs"])
        assert fstats["uncertain_pairs_sent_to_exact"] > 0
        assert fstats["confirmed_pairs"] == ncand  # every sigma > 0
        assert forced == exact_choice
        for (i, k), v in fstats["confirmed"].items():
            assert v == exact_scores[i, k]
