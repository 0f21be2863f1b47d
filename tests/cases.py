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


def _code_values(cs, nominal: np.ndarray, chroma: bool) -> np.ndarray:
    """Nominal luma [0, 1] / chroma [-0.5, 0.5] values -> stored samples."""
    bits = {S.SAMPLE_UINT8: 8, 3: 9, 5: 10, 7: 12, 9: 14, 11: 16}.get(cs.sample)
    if bits is None:  # half / float samples carry full-range values
        v = nominal + (0.5 if chroma else 0.0)
        return v.astype(np.float32 if cs.sample == S.SAMPLE_FLOAT else np.float16)
    if cs.range == S.RANGE_LIMITED:
        code = (128.0 + 224.0 * nominal if chroma else 16.0 + 219.0 * nominal) * (1 << (bits - 8))
    else:
        code = (nominal + (0.5 if chroma else 0.0)) * ((1 << bits) - 1)
    top = (1 << bits) - 1
    code = np.clip(np.round(code), 0, top)
    return code.astype(np.uint8 if bits == 8 else np.uint16)


def _pack(cs, planes, pad):
    out, strides = [], []
    for p, vals in enumerate(planes):
        h, w = cs.plane_shape(p)
        raw = np.ascontiguousarray(vals).view(np.uint8).reshape(h, w * cs.sample_bytes)
        stride = w * cs.sample_bytes + pad * (p + 1)
        buf = np.zeros((h, stride), dtype=np.uint8)
        buf[:, :w * cs.sample_bytes] = raw
        out.append(buf)
        strides.append(stride)
    return out, strides


def perturbed_pair(cs, seed: int, pad: int = 0, noise: float = 0.025):
    """In-gamut random reference planes and a perturbed distorted copy.

    Luma (or R, G, B) nominal values ~ U[0.25, 0.75], chroma ~ U[-0.06, 0.06],
    so every pixel maps to positive linear light; the distorted image adds
    N(0, noise^2) per sample. Rows are padded by pad*(plane+1) bytes.
    """
    rng = np.random.default_rng(seed)
    yuv = cs.colorFamily == S.FAMILY_YUV
    ref_nom, dist_nom = [], []
    for p in range(3):
        h, w = cs.plane_shape(p)
        chroma = yuv and p > 0
        v = rng.uniform(-0.06, 0.06, (h, w)) if chroma else rng.uniform(0.25, 0.75, (h, w))
        ref_nom.append((v, chroma))
        dist_nom.append((v + rng.normal(0, noise, (h, w)), chroma))
    ref, strides = _pack(cs, [_code_values(cs, v, c) for v, c in ref_nom], pad)
    dist, _ = _pack(cs, [_code_values(cs, v, c) for v, c in dist_nom], pad)
    return ref, dist, strides


def smooth_pair(cs, seed: int, pad: int = 0):
    """Natural-like, well-conditioned reference planes and a compressed copy.

    Per plane a sum of 32 random sinusoids (the workload generator's) scaled
    into nominal luma (or R, G, B) [0.2, 0.8] / chroma [-0.06, 0.06]; the
    distorted plane is its 8x8 block DCT quantisation (step 12 on an 8-bit
    scale) plus N(0, 0.002^2). Rows are padded by pad*(plane+1) bytes.
    """
    rng = np.random.default_rng(seed)
    yuv = cs.colorFamily == S.FAMILY_YUV
    ref_nom, dist_nom = [], []
    for p in range(3):
        h, w = cs.plane_shape(p)
        chroma = yuv and p > 0
        f = synth._sinusoid_sum(h, w, rng)
        f = (f - f.min()) / max(f.max() - f.min(), 1e-9) * 2.0 - 1.0
        v = 0.06 * f if chroma else 0.5 + 0.3 * f
        d = synth.dct_quantize(v * 255.0, 12.0) / 255.0 + rng.normal(0, 0.002, (h, w))
        ref_nom.append((v, chroma))
        dist_nom.append((d, chroma))
    ref, strides = _pack(cs, [_code_values(cs, v, c) for v, c in ref_nom], pad)
    dist, _ = _pack(cs, [_code_values(cs, v, c) for v, c in dist_nom], pad)
    return ref, dist, strides


def out_of_gamut_pair(cs, seed: int):
    """Uniform random samples over the whole code range (negative and >1
    linear light included) and a perturbed copy."""
    rng = np.random.default_rng(seed)
    ref, strides = synth.random_planes(cs, rng)
    dist = []
    for p, plane in enumerate(ref):
        d = plane.copy()
        if cs.sample == S.SAMPLE_UINT8:
            d = np.clip(d.astype(np.int32) + rng.integers(-6, 7, d.shape), 0, 255).astype(np.uint8)
        dist.append(d)
    return ref, dist, strides


def wide_range_pair(w=192, h=130, gain=400.0, seed=3):
    """Linear-light float RGB whose right half is `gain` times brighter (linear
    values up to 300, XYB values up to about 7) and a copy perturbed by 2%
    per sample: the blur's fixed first-attempt scale holds centred XYB
    values up to 1.5 only, so items over the bright half and across the edge
    are run again with a scale from their window's bound."""
    cs = S.colorspace(w, h, sample=S.SAMPLE_FLOAT, rng=S.RANGE_FULL, subw=0, subh=0, family=S.FAMILY_RGB,
                      matrix=S.MATRIX_RGB, trc=S.TRC_LINEAR)
    rng = np.random.default_rng(seed)
    ref, dist = [], []
    for _ in range(3):
        v = rng.uniform(0.25, 0.75, (h, w))
        v[:, w // 2:] *= gain
        ref.append(np.ascontiguousarray(v, dtype=np.float32))
        dist.append(np.ascontiguousarray(v * (1.0 + rng.normal(0, 0.02, (h, w))), dtype=np.float32))
    return cs, ref, dist, [w * 4] * 3
