"""FP64 CPU oracle of the SSIMULACRA2 path (test infrastructure only).

ctypes wrapper of oracle/liboracle.so (built from oracle/ssimu2_oracle.c, the
exact formula in FP64, and oracle/ref_float.c, an emulation of the
reference's own FP32 rounding, by oracle/Makefile). Only tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py use it, as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(_HERE, "_ref", "libvship_ref.so")

_lib = None
_D = ctypes.POINTER(ctypes.c_double)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        from vship_amd.ssimu2 import Colorspace
        CS = ctypes.POINTER(Colorspace)
        vp3 = ctypes.c_void_p * 3
        i64 = ctypes.POINTER(ctypes.c_int64)
        L.oracle_weight.restype = ctypes.c_double
        L.oracle_weight.argtypes = [ctypes.c_int]
        L.oracle_map_active.argtypes = [ctypes.c_int] * 3
        L.oracle_range_to_full.restype = ctypes.c_double
        L.oracle_range_to_full.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_catmull_rom.restype = ctypes.c_double
        L.oracle_catmull_rom.argtypes = [ctypes.c_double] * 5
        L.oracle_upsample_plane.argtypes = [_D, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, _D]
        L.oracle_transfer.restype = ctypes.c_double
        L.oracle_transfer.argtypes = [ctypes.c_double, ctypes.c_int]
        L.oracle_yuv_to_linear.argtypes = [_D, ctypes.c_int, ctypes.c_int]
        L.oracle_primaries_to_709.argtypes = [_D, ctypes.c_int]
        L.oracle_final_dims.argtypes = [CS, i64, i64]
        L.oracle_linear_rgb.argtypes = [CS, vp3, i64, _D]
        L.oracle_downsample.argtypes = [_D, ctypes.c_int64, ctypes.c_int64, _D]
        L.oracle_xyb.argtypes = [_D, _D, _D, ctypes.c_int64]
        L.oracle_total_scale_size.restype = ctypes.c_int64
        L.oracle_total_scale_size.argtypes = [ctypes.c_int64, ctypes.c_int64]
        L.oracle_xyb_pyramid.argtypes = [_D, ctypes.c_int64, ctypes.c_int64, _D]
        L.oracle_gauss_taps.argtypes = [_D]
        L.oracle_blur.argtypes = [_D, ctypes.c_int64, ctypes.c_int64, _D]
        L.oracle_maps.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_int64, _D, _D, _D]
        L.oracle_plane_features.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_int64, _D]
        L.oracle_final_score.restype = ctypes.c_double
        L.oracle_final_score.argtypes = [_D]
        L.oracle_features_from_linear.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _D]
        L.oracle_ssimu2.argtypes = [CS, vp3, i64, CS, vp3, i64, ctypes.c_int, _D, _D]
        L.oracle_reference_float.argtypes = [CS, vp3, i64, CS, vp3, i64, _D, _D]
        _F = ctypes.POINTER(ctypes.c_float)
        L.oracle_reference_float_from_linear.restype = ctypes.c_double
        L.oracle_reference_float_from_linear.argtypes = [_F, _F, ctypes.c_int64, ctypes.c_int64, _D]
        L.oracle_reference_float_perturbed.restype = ctypes.c_double
        L.oracle_reference_float_perturbed.argtypes = [_F, _F, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64,
                                                       ctypes.c_double]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def weights() -> np.ndarray:
    return np.array([lib().oracle_weight(i) for i in range(108)])


def map_active(plane: int, scale: int, m: int) -> bool:
    return bool(lib().oracle_map_active(plane, scale, m))


def gauss_taps() -> np.ndarray:
    out = np.zeros(17)
    lib().oracle_gauss_taps(_dp(out))
    return out


def blur(img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.float64)
    out = np.zeros_like(img)
    lib().oracle_blur(_dp(img), img.shape[1], img.shape[0], _dp(out))
    return out


def downsample(img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.float64)
    h, w = img.shape
    out = np.zeros(((h - 1) // 2 + 1, (w - 1) // 2 + 1))
    lib().oracle_downsample(_dp(img), w, h, _dp(out))
    return out


def xyb(r, g, b):
    r, g, b = (np.array(v, dtype=np.float64, copy=True).reshape(-1) for v in (r, g, b))
    lib().oracle_xyb(_dp(r), _dp(g), _dp(b), r.size)
    return r, g, b


def transfer(v: float, trc: int) -> float:
    return lib().oracle_transfer(float(v), trc)


def range_to_full(v: float, sample: int, rng: int, chroma: bool) -> float:
    return lib().oracle_range_to_full(float(v), sample, rng, int(chroma))


def yuv_to_linear(y, u, v, matrix: int, trc: int):
    a = np.array([y, u, v], dtype=np.float64)
    lib().oracle_yuv_to_linear(_dp(a), matrix, trc)
    return a


def primaries_to_709(rgb, prim: int):
    a = np.array(rgb, dtype=np.float64)
    lib().oracle_primaries_to_709(_dp(a), prim)
    return a


def catmull_rom(pm1, p0, p1, p2, t) -> float:
    return lib().oracle_catmull_rom(pm1, p0, p1, p2, t)


def upsample_plane(small: np.ndarray, w: int, h: int, subw: int, subh: int, loc: int) -> np.ndarray:
    small = np.ascontiguousarray(small, dtype=np.float64)
    out = np.zeros((h, w))
    rc = lib().oracle_upsample_plane(_dp(small), w, h, subw, subh, loc, _dp(out))
    if rc != 0:
        raise ValueError("unsupported subsampling")
    return out


def maps(a: np.ndarray, b: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    s, ar, d = np.zeros_like(a), np.zeros_like(a), np.zeros_like(a)
    lib().oracle_maps(_dp(a), _dp(b), a.shape[1], a.shape[0], _dp(s), _dp(ar), _dp(d))
    return s, ar, d


def final_score(features) -> float:
    f = np.ascontiguousarray(features, dtype=np.float64)
    return lib().oracle_final_score(_dp(f))


def _planes(planes, strides):
    vp = (ctypes.c_void_p * 3)(*[p.ctypes.data for p in planes])
    ls = (ctypes.c_int64 * 3)(*[int(s) for s in strides])
    return vp, ls


def linear_rgb(cs, planes, strides) -> np.ndarray:
    w, h = ctypes.c_int64(), ctypes.c_int64()
    lib().oracle_final_dims(ctypes.byref(cs), ctypes.byref(w), ctypes.byref(h))
    out = np.zeros((3, h.value, w.value))
    vp, ls = _planes(planes, strides)
    rc = lib().oracle_linear_rgb(ctypes.byref(cs), vp, ls, _dp(out))
    if rc != 0:
        raise ValueError(f"oracle_linear_rgb failed ({rc})")
    return out


def xyb_pyramid(rgb: np.ndarray):
    """List over scales 0..5 of (3, h_s, w_s) XYB arrays."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float64)
    _, h, w = rgb.shape
    T = lib().oracle_total_scale_size(w, h)
    out = np.zeros(3 * T)
    lib().oracle_xyb_pyramid(_dp(rgb), w, h, _dp(out))
    out = out.reshape(3, T)
    levels, off, cw, ch = [], 0, w, h
    for _ in range(6):
        levels.append(out[:, off:off + cw * ch].reshape(3, ch, cw).copy())
        off += cw * ch
        cw, ch = (cw - 1) // 2 + 1, (ch - 1) // 2 + 1
    return levels


def ssimu2(cs1, planes1, strides1, cs2, planes2, strides2, prune: bool = True):
    """(score, features[108]) of the whole path on raw planes."""
    f = np.zeros(108)
    s = ctypes.c_double()
    p1, l1 = _planes(planes1, strides1)
    p2, l2 = _planes(planes2, strides2)
    rc = lib().oracle_ssimu2(ctypes.byref(cs1), p1, l1, ctypes.byref(cs2), p2, l2, int(prune), _dp(f),
                             ctypes.byref(s))
    if rc != 0:
        raise ValueError(f"oracle_ssimu2 failed ({rc})")
    return s.value, f


def features_from_linear(rgb1: np.ndarray, rgb2: np.ndarray, prune: bool = True) -> np.ndarray:
    rgb1 = np.ascontiguousarray(rgb1, dtype=np.float64)
    rgb2 = np.ascontiguousarray(rgb2, dtype=np.float64)
    f = np.zeros(108)
    rc = lib().oracle_features_from_linear(_dp(rgb1), _dp(rgb2), rgb1.shape[2], rgb1.shape[1], int(prune), _dp(f))
    if rc != 0:
        raise ValueError("oracle_features_from_linear failed")
    return f


def reference_float(cs1, planes1, strides1, cs2, planes2, strides2):
    """(score, features[108]) of the emulation of the reference's FP32 arithmetic (oracle/ref_float.c):
    what the reference libvship returns up to the float rounding of its colour front-end."""
    f = np.zeros(108)
    s = ctypes.c_double()
    p1, l1 = _planes(planes1, strides1)
    p2, l2 = _planes(planes2, strides2)
    rc = lib().oracle_reference_float(ctypes.byref(cs1), p1, l1, ctypes.byref(cs2), p2, l2, _dp(f), ctypes.byref(s))
    if rc != 0:
        raise ValueError(f"oracle_reference_float failed ({rc})")
    return s.value, f


def reference_float_from_linear(rgb1: np.ndarray, rgb2: np.ndarray):
    """The emulated reference score of two (3, h, w) linear-RGB float32 images."""
    a = np.ascontiguousarray(rgb1, dtype=np.float32)
    b = np.ascontiguousarray(rgb2, dtype=np.float32)
    f = np.zeros(108)
    fp = ctypes.POINTER(ctypes.c_float)
    s = lib().oracle_reference_float_from_linear(a.ctypes.data_as(fp), b.ctypes.data_as(fp), a.shape[2], a.shape[1],
                                                 _dp(f))
    return s, f


def reference_float_spread(rgb1: np.ndarray, rgb2: np.ndarray, draws: int = 16, frac: float = 0.5):
    """Emulated reference scores with a fraction `frac` of the XYB values moved by one ulp, one per seed:
    the spread is how far last-bit differences of the reference's libm can move its score."""
    a = np.ascontiguousarray(rgb1, dtype=np.float32)
    b = np.ascontiguousarray(rgb2, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    return np.array([lib().oracle_reference_float_perturbed(a.ctypes.data_as(fp), b.ctypes.data_as(fp), a.shape[2],
                                                            a.shape[1], 1000 + k, frac) for k in range(draws)])
