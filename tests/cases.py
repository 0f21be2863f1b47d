"""Colourspace cases and seeded inputs shared by the parity tests."""
from __future__ import annotations

import numpy as np

from vship_amd import ssimu2 as S
from vship_amd import synth

# (name, kwargs for ssimu2.colorspace)
COLOR_CASES = [
    ("yuv420p8_709_limited_left", dict()),
    ("yuv420p10_709_limited", dict(sample=S.SAMPLE_UINT10)),
    ("yuv420p8_709_full_center", dict(rng=S.RANGE_FULL, loc=S.CHROMA_CENTER)),
    ("yuv420p8_topleft", dict(loc=S.CHROMA_TOPLEFT)),
    ("yuv420p8_top", dict(loc=S.CHROMA_TOP)),
    ("yuv422p8", dict(subw=1, subh=0)),
    ("yuv440p8", dict(subw=0, subh=1)),
    ("yuv411p8_left", dict(subw=2, subh=0)),
    ("yuv411p8_center", dict(subw=2, subh=0, loc=S.CHROMA_CENTER)),
    ("yuv410p8", dict(subw=2, subh=1)),
    ("yuv444p16_2020ncl_pq", dict(sample=S.SAMPLE_UINT16, subw=0, subh=0, matrix=S.MATRIX_BT2020_NCL,
                                  trc=S.TRC_PQ, prim=S.PRIM_BT2020)),
    ("yuv420p12_2020cl_hlg", dict(sample=S.SAMPLE_UINT12, matrix=S.MATRIX_BT2020_CL, trc=S.TRC_HLG,
                                  prim=S.PRIM_BT2020)),
    ("yuv420p10_ictcp_pq", dict(sample=S.SAMPLE_UINT10, matrix=S.MATRIX_BT2100_ICTCP, trc=S.TRC_PQ,
                                prim=S.PRIM_BT2020)),
    ("yuv420p8_470bg_601", dict(matrix=S.MATRIX_BT470_BG, trc=S.TRC_BT601, prim=S.PRIM_BT470_BG)),
    ("yuv420p9_170m_470m", dict(sample=S.SAMPLE_UINT9, matrix=S.MATRIX_ST170_M, trc=S.TRC_BT470_M,
                                prim=S.PRIM_ST170_M)),
    ("yuv420p14_st240", dict(sample=S.SAMPLE_UINT14, trc=S.TRC_ST240_M, prim=S.PRIM_ST240_M)),
    ("yuv444p8_st428_xyz", dict(subw=0, subh=0, trc=S.TRC_ST428, prim=S.PRIM_INTERNAL)),
    ("rgb8_srgb_full", dict(rng=S.RANGE_FULL, subw=0, subh=0, family=S.FAMILY_RGB, matrix=S.MATRIX_RGB,
                            trc=S.TRC_SRGB)),
    ("rgbs_linear", dict(sample=S.SAMPLE_FLOAT, rng=S.RANGE_FULL, subw=0, subh=0, family=S.FAMILY_RGB,
                         matrix=S.MATRIX_RGB, trc=S.TRC_LINEAR)),
    ("rgbh_709", dict(sample=S.SAMPLE_HALF, rng=S.RANGE_FULL, subw=0, subh=0, family=S.FAMILY_RGB,
                      matrix=S.MATRIX_RGB, trc=S.TRC_BT709, prim=S.PRIM_BT470_M)),
]


def perturbed_pair(cs, seed: int, pad: int = 0):
    """Random reference planes and a mildly perturbed distorted copy."""
    rng = np.random.default_rng(seed)
    ref, strides = synth.random_planes(cs, rng, pad=pad)
    dist = []
    nb = cs.sample_bytes
    for p, (plane, stride) in enumerate(zip(ref, strides)):
        h, w = cs.plane_shape(p)
        d = plane.copy()
        body = d[:, :w * nb]
        if cs.sample == S.SAMPLE_UINT8:
            v = body.astype(np.int32) + rng.integers(-6, 7, body.shape)
            d[:, :w * nb] = np.clip(v, 0, 255).astype(np.uint8)
        elif cs.sample == S.SAMPLE_FLOAT:
            v = body.view(np.float32) + rng.normal(0, 0.02, (h, w)).astype(np.float32)
            d[:, :w * nb] = np.clip(v, 0, 1).astype(np.float32).view(np.uint8)
        elif cs.sample == S.SAMPLE_HALF:
            v = body.view(np.float16).astype(np.float32) + rng.normal(0, 0.02, (h, w)).astype(np.float32)
            d[:, :w * nb] = np.clip(v, 0, 1).astype(np.float16).view(np.uint8)
        else:
            bits = {3: 9, 5: 10, 7: 12, 9: 14, 11: 16}[cs.sample]
            v = body.view(np.uint16).astype(np.int64) + rng.integers(-6, 7, (h, w)) * (1 << (bits - 8))
            d[:, :w * nb] = np.clip(v, 0, (1 << bits) - 1).astype(np.uint16).view(np.uint8)
        dist.append(d)
    return ref, dist, strides
