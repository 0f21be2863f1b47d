"""The seeded inputs behind tests/golden/ref_scores.json (regenerated
bit-identically by numpy PCG64 / scipy on any machine)."""
from __future__ import annotations

from vship_amd import ssimu2 as S
from vship_amd import synth
from tests import cases


def golden_inputs():
    """Yields (name, cs1, planes1, strides1, cs2, planes2, strides2)."""
    ref, dist = synth.srgb_pair(256, 256, seed=0)
    cs = S.srgb_rgb_full(256, 256)
    yield "srgb256_seed0", cs, ref, [256] * 3, cs, dist, [256] * 3
    yield "srgb256_seed0_identity", cs, ref, [256] * 3, cs, ref, [256] * 3
    for w, h, frames in ((1920, 1080, (0, 1, 2)), (1280, 720, (0, 5)), (640, 360, (0, 3))):
        seq = synth.YuvSequence(w, h, frames=max(frames) + 1, seed=3)
        c = S.colorspace(w, h)
        for i in frames:
            r, d, _ = seq.frame(i)
            st = S.host_strides(r)
            yield f"yuv420p8_{w}x{h}_seed3_frame{i}", c, r, st, c, d, st
    seq = synth.YuvSequence(1920, 1080, frames=1, bits=10, seed=4)
    c = S.colorspace(1920, 1080, sample=S.SAMPLE_UINT10)
    r, d, _ = seq.frame(0)
    st = S.host_strides(r)
    yield "yuv420p10_1920x1080_seed4_frame0", c, r, st, c, d, st
    for name, kw in cases.COLOR_CASES:
        c = S.colorspace(131, 75, **kw)
        r, d, st = cases.perturbed_pair(c, seed=3, pad=5)
        yield f"random131x75_{name}", c, r, st, c, d, st
    # the 16 frames of the throughput comparison with the reference
    seq = synth.YuvSequence(1920, 1080, frames=16, seed=7)
    c = S.colorspace(1920, 1080)
    for i in range(16):
        r, d, _ = seq.frame(i)
        st = S.host_strides(r)
        yield f"yuv420p8_1920x1080_seed7_frame{i}", c, r, st, c, d, st
    # every colourspace on smooth (well-conditioned) content
    for name, kw in cases.COLOR_CASES:
        c = S.colorspace(320, 180, **kw)
        r, d, st = cases.smooth_pair(c, seed=5, pad=5)
        yield f"smooth320x180_{name}", c, r, st, c, d, st
    # the ST240 case's ingredients one at a time, on its random content
    for name, kw, pad in (("yuv420p14_st240_dense", dict(sample=S.SAMPLE_UINT14, trc=S.TRC_ST240_M,
                                                           prim=S.PRIM_ST240_M), 0),
                          ("yuv420p14_709", dict(sample=S.SAMPLE_UINT14), 5),
                          ("yuv420p8_st240trc", dict(trc=S.TRC_ST240_M), 5),
                          ("yuv420p8_st240prim", dict(prim=S.PRIM_ST240_M), 5)):
        c = S.colorspace(131, 75, **kw)
        r, d, st = cases.perturbed_pair(c, seed=3, pad=pad)
        yield f"random131x75_{name}", c, r, st, c, d, st
