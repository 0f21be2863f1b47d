"""Identity stress test on the GPU, after the reference's only correctness test
(test/libvshipTestIdentity.cpp:214-251,285-346,389-461): for every combination
of sample type (7), range (2), chroma subsampling (6), chroma siting (4) and
crop (3), at 720x480 and 1920x1080, at native size and resized to 720x480, a
pair of identical images whose row padding bytes differ scores exactly 100.

The reference's fixed colour description is kept (YUV family, RGB matrix,
BT.709 transfer, internal primaries) and so are its line sizes (samples of a
row, then 47, 38 and 97 padding bytes for planes 0, 1, 2; odd paddings leave
16-bit samples unaligned). Samples are drawn over each type's whole range (the
reference's generator reaches only 0..3 for the 9- to 16-bit types, :280).
One image per size and sample type serves all of its cases; the padding bytes
of both images are redrawn for every case."""
import itertools

import numpy as np
import pytest

from vship_amd import ssimu2 as S

pytestmark = pytest.mark.gpu

SAMPLES = [S.SAMPLE_FLOAT, S.SAMPLE_UINT16, S.SAMPLE_UINT14, S.SAMPLE_UINT12, S.SAMPLE_UINT10, S.SAMPLE_UINT9,
           S.SAMPLE_UINT8]
BITS = {S.SAMPLE_UINT16: 16, S.SAMPLE_UINT14: 14, S.SAMPLE_UINT12: 12, S.SAMPLE_UINT10: 10, S.SAMPLE_UINT9: 9,
        S.SAMPLE_UINT8: 8}
RANGES = [S.RANGE_LIMITED, S.RANGE_FULL]
SUBSAMPLINGS = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]
SITINGS = [S.CHROMA_LEFT, S.CHROMA_CENTER, S.CHROMA_TOP, S.CHROMA_TOPLEFT]
CROPS = [(0, 0, 0, 0), (10, 10, 10, 10), (1, 3, 5, 7)]
PADDING = (47, 38, 97)


def _samples(rng, sample, n):
    """n random samples of the type as little-endian bytes."""
    if sample == S.SAMPLE_FLOAT:
        return rng.random(n, dtype=np.float32).view(np.uint8)
    if sample == S.SAMPLE_UINT8:
        return rng.integers(0, 256, n, dtype=np.uint8)
    return rng.integers(0, 1 << BITS[sample], n, dtype=np.uint16).astype("<u2").view(np.uint8)


class Content:
    """Random sample rows of one size and sample type (the largest chroma
    planes any subsampling needs), cut to each case's plane shapes."""

    def __init__(self, w, h, sample, seed):
        rng = np.random.default_rng(seed)
        self.bytes = S.SAMPLE_BYTES[sample]
        self.planes = [_samples(rng, sample, w * h).reshape(h, w * self.bytes) for _ in range(3)]


def _image(content, w, h, subw, subh, rng):
    """Planes with the reference's line sizes; padding bytes drawn from rng."""
    b = content.bytes
    row = b * w
    planes, strides = [], []
    for p in range(3):
        pw, ph = (w, h) if p == 0 else (w >> subw, h >> subh)
        ls = (row if p == 0 else row >> subw) + PADDING[p]
        buf = np.empty((ph, ls), dtype=np.uint8)
        buf[:, :b * pw] = content.planes[p][:ph, :b * pw]
        buf[:, b * pw:] = rng.integers(0, 256, (ph, ls - b * pw), dtype=np.uint8)
        planes.append(buf.reshape(-1))
        strides.append(ls)
    return planes, strides


@pytest.mark.parametrize("size", [(720, 480), (1920, 1080)], ids=["720x480", "1920x1080"])
@pytest.mark.parametrize("target", [(-1, -1), (720, 480)], ids=["native", "to720x480"])
def test_identity_stress(size, target):
    w, h = size
    rng = np.random.default_rng(w * 7 + target[0])
    failures, n = [], 0
    for si, sample in enumerate(SAMPLES):
        content = Content(w, h, sample, seed=1000 * si + w)
        for rg, (subw, subh), loc, crop in itertools.product(RANGES, SUBSAMPLINGS, SITINGS, CROPS):
            cs = S.colorspace(w, h, sample=sample, rng=rg, subw=subw, subh=subh, loc=loc, family=S.FAMILY_YUV,
                              matrix=S.MATRIX_RGB, trc=S.TRC_BT709, prim=S.PRIM_INTERNAL)
            cs.target_width, cs.target_height = target
            cs.crop = S.CropRectangle(*crop)
            im1, ls1 = _image(content, w, h, subw, subh, rng)
            im2, ls2 = _image(content, w, h, subw, subh, rng)
            with S.SSIMU2Handler(cs, cs) as hd:
                score = hd.compute(im1, im2, ls1, ls2)
            n += 1
            if score != 100.0:
                failures.append((sample, rg, subw, subh, loc, crop, score))
    print(f"identity {w}x{h} target {target}: {n} cases, {len(failures)} not 100", flush=True)
    assert n == 7 * 2 * 6 * 4 * 3
    assert not failures, failures[:10]
