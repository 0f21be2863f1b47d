"""Python host mirror of the SSIMULACRA2 C ABI of libvship_amd.so.

The names and semantics follow the reference's C API (src/VshipAPI.h:98-119,
src/VshipColor.h:7-96): a handler is created from two colourspace
descriptors, computes one score per frame pair and is freed explicitly.
Errors surface as :class:`VshipError` carrying the Vship_Exception code and
the detailed message (src/VshipLib.cpp:151-167,262-286).

The library is the HIP build in this package directory; there is no CPU
fallback: a missing library raises at load time.
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvship_amd.so")

# ---------------------------------------------------------------- enums

SAMPLE_FLOAT, SAMPLE_HALF, SAMPLE_UINT8, SAMPLE_UINT9 = 0, 1, 2, 3
SAMPLE_UINT10, SAMPLE_UINT12, SAMPLE_UINT14, SAMPLE_UINT16 = 5, 7, 9, 11
RANGE_LIMITED, RANGE_FULL = 0, 1
CHROMA_LEFT, CHROMA_CENTER, CHROMA_TOPLEFT, CHROMA_TOP = 0, 1, 2, 3
FAMILY_YUV, FAMILY_RGB = 0, 1
MATRIX_RGB, MATRIX_BT709, MATRIX_BT470_BG, MATRIX_ST170_M = 0, 1, 5, 6
MATRIX_BT2020_NCL, MATRIX_BT2020_CL, MATRIX_BT2100_ICTCP = 9, 10, 14
TRC_BT709, TRC_BT470_M, TRC_BT470_BG, TRC_BT601, TRC_ST240_M = 1, 4, 5, 6, 7
TRC_LINEAR, TRC_SRGB, TRC_PQ, TRC_ST428, TRC_HLG = 8, 13, 16, 17, 18
PRIM_INTERNAL, PRIM_BT709, PRIM_BT470_M, PRIM_BT470_BG = -1, 1, 4, 5
PRIM_ST170_M, PRIM_ST240_M, PRIM_BT2020 = 6, 7, 9

SAMPLE_BYTES = {SAMPLE_FLOAT: 4, SAMPLE_HALF: 2, SAMPLE_UINT8: 1, SAMPLE_UINT9: 2,
                SAMPLE_UINT10: 2, SAMPLE_UINT12: 2, SAMPLE_UINT14: 2, SAMPLE_UINT16: 2}

ERROR_NAMES = {
    0: "NoError", 1: "OutOfVRAM", 2: "OutOfRAM", 12: "HIPError", 3: "BadDisplayModel",
    4: "DifferingInputType", 5: "NonRGBSInput", 13: "BadPath", 14: "BadJson",
    6: "DeviceCountError", 7: "NoDeviceDetected", 8: "BadDeviceArgument", 9: "BadDeviceCode",
    10: "BadHandler", 11: "BadPointer", 15: "BadErrorType",
}


class ChromaSubsample(ctypes.Structure):
    _fields_ = [("subw", ctypes.c_int), ("subh", ctypes.c_int)]


class CropRectangle(ctypes.Structure):
    _fields_ = [("top", ctypes.c_int), ("bottom", ctypes.c_int), ("left", ctypes.c_int), ("right", ctypes.c_int)]


class Colorspace(ctypes.Structure):
    """Vship_Colorspace_t (reference src/VshipColor.h:81-96), passed by value."""
    _fields_ = [
        ("width", ctypes.c_int64), ("height", ctypes.c_int64),
        ("target_width", ctypes.c_int64), ("target_height", ctypes.c_int64),
        ("sample", ctypes.c_int), ("range", ctypes.c_int),
        ("subsampling", ChromaSubsample), ("chromaLocation", ctypes.c_int),
        ("colorFamily", ctypes.c_int), ("YUVMatrix", ctypes.c_int),
        ("transferFunction", ctypes.c_int), ("primaries", ctypes.c_int),
        ("crop", CropRectangle),
    ]

    def plane_shape(self, p: int):
        """(height, width) of plane p before chroma upsampling."""
        if p == 0:
            return int(self.height), int(self.width)
        sw, sh = self.subsampling.subw, self.subsampling.subh
        return ((int(self.height) - 1) >> sh) + 1, ((int(self.width) - 1) >> sw) + 1

    @property
    def sample_bytes(self) -> int:
        return SAMPLE_BYTES[self.sample]


def colorspace(width, height, sample=SAMPLE_UINT8, rng=RANGE_LIMITED, subw=1, subh=1, loc=CHROMA_LEFT,
               family=FAMILY_YUV, matrix=MATRIX_BT709, trc=TRC_BT709, prim=PRIM_BT709) -> Colorspace:
    """Build a descriptor; defaults are 8-bit YUV 4:2:0 BT.709 limited."""
    cs = Colorspace()
    cs.width, cs.height, cs.target_width, cs.target_height = width, height, -1, -1
    cs.sample, cs.range = sample, rng
    cs.subsampling = ChromaSubsample(subw, subh)
    cs.chromaLocation, cs.colorFamily, cs.YUVMatrix = loc, family, matrix
    cs.transferFunction, cs.primaries = trc, prim
    cs.crop = CropRectangle(0, 0, 0, 0)
    return cs


def srgb_rgb_full(width, height) -> Colorspace:
    """8-bit planar RGB, full range, sRGB transfer, BT.709 primaries."""
    return colorspace(width, height, SAMPLE_UINT8, RANGE_FULL, 0, 0, CHROMA_LEFT, FAMILY_RGB,
                      MATRIX_RGB, TRC_SRGB, PRIM_BT709)


class Handler(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint)]


class BatchTuning(ctypes.Structure):
    """VshipAMD_BatchTuning: blur tile shape id, blur and front-end WGs per CU,
    generic_front (1: no layout instantiation of the front-end)."""
    _fields_ = [("blur_config", ctypes.c_int), ("blur_waves", ctypes.c_int), ("front_waves", ctypes.c_int),
                ("generic_front", ctypes.c_int)]

    def as_dict(self):
        return {"blur_config": self.blur_config, "blur_waves": self.blur_waves, "front_waves": self.front_waves,
                "generic_front": self.generic_front}


class Version(ctypes.Structure):
    _fields_ = [("major", ctypes.c_int), ("minor", ctypes.c_int), ("minorMinor", ctypes.c_int),
                ("backend", ctypes.c_int)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("VRAMSize", ctypes.c_uint64), ("integrated", ctypes.c_int),
                ("MultiProcessorCount", ctypes.c_int), ("WarpSize", ctypes.c_int)]


class VshipError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)} ({code}): {message}")
        self.code = code
        self.message = message


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libvship_amd.so (built by __graft_entry__.build()). Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build the HIP library first (python -c "
                           "'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    u8pp = P(ctypes.c_void_p)
    i64p = P(ctypes.c_int64)
    E = ctypes.c_int
    sigs = {
        "Vship_GetVersion": (Version, []),
        "Vship_GetDeviceCount": (E, [P(ctypes.c_int)]),
        "Vship_SetDevice": (E, [ctypes.c_int]),
        "Vship_GetDeviceInfo": (E, [P(DeviceInfo), ctypes.c_int]),
        "Vship_GPUFullCheck": (E, [ctypes.c_int]),
        "Vship_GetErrorMessage": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
        "Vship_GetDetailedLastError": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
        "Vship_PinnedMalloc": (E, [P(ctypes.c_void_p), ctypes.c_uint64]),
        "Vship_PinnedFree": (E, [ctypes.c_void_p]),
        "Vship_SSIMU2Init": (E, [P(Handler), Colorspace, Colorspace]),
        "Vship_SSIMU2Init2": (E, [P(Handler), Colorspace, Colorspace, ctypes.c_int]),
        "Vship_SSIMU2Free": (E, [Handler]),
        "Vship_ComputeSSIMU2": (E, [Handler, P(ctypes.c_double), u8pp, u8pp, i64p, i64p]),
        "Vship_SSIMU2GetDetailedLastError": (ctypes.c_int, [Handler, ctypes.c_char_p, ctypes.c_int]),
        "VshipAMD_ComputeSSIMU2Device": (E, [Handler, P(ctypes.c_double), u8pp, u8pp, i64p, i64p]),
        "VshipAMD_ComputeSSIMU2DeviceNoAlloc": (E, [Colorspace, Colorspace, P(ctypes.c_double), u8pp, u8pp, i64p,
                                                    i64p, ctypes.c_int]),
        "VshipAMD_SSIMU2Features": (E, [Handler, P(ctypes.c_double)]),
        "VshipAMD_SSIMU2DebugPlane": (E, [Handler, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_float),
                                          i64p, i64p]),
        "VshipAMD_SSIMU2Profile": (E, [Handler, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double),
                                       P(ctypes.c_double)]),
        "VshipAMD_SSIMU2Work": (E, [Handler, P(ctypes.c_double), P(ctypes.c_double)]),
        "VshipAMD_SSIMU2BatchInit": (E, [P(Handler), ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "VshipAMD_SSIMU2BatchFree": (E, [Handler]),
        "VshipAMD_SSIMU2BatchCompute": (E, [Handler, ctypes.c_int, P(Colorspace), P(Colorspace), u8pp, u8pp, i64p,
                                            i64p, P(ctypes.c_double)]),
        "VshipAMD_SSIMU2BatchComputeScales": (E, [Handler, ctypes.c_int, P(Colorspace), P(Colorspace), u8pp, u8pp,
                                                  i64p, i64p, ctypes.c_int, P(ctypes.c_double)]),
        "VshipAMD_SSIMU2BatchCompleteScales": (E, [Handler, ctypes.c_int, P(ctypes.c_int), ctypes.c_int,
                                                   P(ctypes.c_double)]),
        "VshipAMD_SSIMU2BatchAutotune": (E, [Handler, ctypes.c_int, P(Colorspace), P(Colorspace), u8pp, u8pp, i64p,
                                             i64p, ctypes.c_int, P(BatchTuning), P(ctypes.c_double)]),
        "VshipAMD_SSIMU2BatchSetTuning": (E, [Handler, BatchTuning]),
        "VshipAMD_SSIMU2BatchGetTuning": (E, [Handler, P(BatchTuning)]),
        "VshipAMD_SSIMU2BatchFeatures": (E, [Handler, P(ctypes.c_double)]),
        "VshipAMD_SSIMU2BatchProfile": (E, [Handler, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double),
                                            P(ctypes.c_double), P(ctypes.c_double)]),
        "VshipAMD_SSIMU2BatchWork": (E, [Handler, P(ctypes.c_double), P(ctypes.c_double)]),
        "VshipAMD_DeviceMalloc": (E, [P(ctypes.c_void_p), ctypes.c_uint64, ctypes.c_int]),
        "VshipAMD_DeviceFree": (E, [ctypes.c_void_p]),
        "VshipAMD_MemcpyHtoD": (E, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
        "VshipAMD_MemcpyDtoH": (E, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
        "VshipAMD_DeviceSynchronize": (E, []),
        "VshipAMD_DeviceCopy": (E, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                    P(ctypes.c_double)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _detail() -> str:
    lib = load_library()
    n = lib.Vship_GetDetailedLastError(None, 0)
    buf = ctypes.create_string_buffer(n)
    lib.Vship_GetDetailedLastError(buf, n)
    return buf.value.decode(errors="replace")


def check(code: int) -> None:
    if code != 0:
        raise VshipError(code, _detail())


def device_count() -> int:
    n = ctypes.c_int(0)
    check(load_library().Vship_GetDeviceCount(ctypes.byref(n)))
    return n.value


def error_message(code: int) -> str:
    lib = load_library()
    n = lib.Vship_GetErrorMessage(code, None, 0)
    buf = ctypes.create_string_buffer(n)
    lib.Vship_GetErrorMessage(code, buf, n)
    return buf.value.decode()


def _plane_args(planes: Sequence, strides: Sequence[int]):
    ptrs = (ctypes.c_void_p * 3)()
    ls = (ctypes.c_int64 * 3)()
    for i in range(3):
        p = planes[i]
        if isinstance(p, np.ndarray):
            if not p.flags["C_CONTIGUOUS"]:
                raise ValueError("host planes must be C-contiguous buffers")
            ptrs[i] = p.ctypes.data
        else:
            ptrs[i] = int(p)
        ls[i] = int(strides[i])
    return ptrs, ls


def host_strides(planes: Sequence[np.ndarray]):
    """Line size in bytes of 2-D (or padded 2-D) numpy planes."""
    return [int(p.strides[0]) for p in planes]


class DeviceBuffer:
    """Raw device allocation through the library (no other HIP binding needed)."""

    def __init__(self, nbytes: int, gpu_id: int = 0):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        check(load_library().VshipAMD_DeviceMalloc(ctypes.byref(self.ptr), self.nbytes, gpu_id))

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes
        check(load_library().VshipAMD_MemcpyHtoD(ctypes.c_void_p(self.address + offset), arr.ctypes.data, arr.nbytes))

    def download(self, nbytes: int, offset: int = 0) -> np.ndarray:
        out = np.empty(nbytes, dtype=np.uint8)
        check(load_library().VshipAMD_MemcpyDtoH(out.ctypes.data, ctypes.c_void_p(self.address + offset), nbytes))
        return out

    def free(self) -> None:
        if self.ptr.value:
            load_library().VshipAMD_DeviceFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SSIMU2Handler:
    """One Vship_SSIMU2Handler: Init2 on construction, Free on close()."""

    def __init__(self, src: Colorspace, dis: Colorspace, gpu_id: int = 0):
        self._lib = load_library()
        self.src, self.dis = src, dis
        self.handle = Handler()
        code = self._lib.Vship_SSIMU2Init2(ctypes.byref(self.handle), src, dis, gpu_id)
        if code != 0:
            msg = _detail()
            self._lib.Vship_SSIMU2Free(self.handle)
            raise VshipError(code, msg)
        self._open = True

    def _check(self, code: int) -> None:
        if code != 0:
            n = self._lib.Vship_SSIMU2GetDetailedLastError(self.handle, None, 0)
            buf = ctypes.create_string_buffer(max(n, 1))
            self._lib.Vship_SSIMU2GetDetailedLastError(self.handle, buf, max(n, 1))
            raise VshipError(code, buf.value.decode(errors="replace"))

    def compute(self, planes1: Sequence[np.ndarray], planes2: Sequence[np.ndarray],
                strides1: Sequence[int] | None = None, strides2: Sequence[int] | None = None) -> float:
        """Vship_ComputeSSIMU2 on host planes (numpy arrays, rows may be padded)."""
        s1 = strides1 if strides1 is not None else host_strides(planes1)
        s2 = strides2 if strides2 is not None else host_strides(planes2)
        p1, l1 = _plane_args(planes1, s1)
        p2, l2 = _plane_args(planes2, s2)
        score = ctypes.c_double(0.0)
        self._check(self._lib.Vship_ComputeSSIMU2(self.handle, ctypes.byref(score), p1, p2, l1, l2))
        return score.value

    def compute_device(self, dptrs1: Sequence[int], dptrs2: Sequence[int], strides1: Sequence[int],
                       strides2: Sequence[int]) -> float:
        """VshipAMD_ComputeSSIMU2Device on device-resident planes."""
        p1, l1 = _plane_args(dptrs1, strides1)
        p2, l2 = _plane_args(dptrs2, strides2)
        score = ctypes.c_double(0.0)
        self._check(self._lib.VshipAMD_ComputeSSIMU2Device(self.handle, ctypes.byref(score), p1, p2, l1, l2))
        return score.value

    def features(self) -> np.ndarray:
        out = np.zeros(108, dtype=np.float64)
        self._check(self._lib.VshipAMD_SSIMU2Features(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def debug_plane(self, image: int, plane: int, scale: int) -> np.ndarray:
        w, h = ctypes.c_int64(0), ctypes.c_int64(0)
        self._check(self._lib.VshipAMD_SSIMU2DebugPlane(self.handle, image, plane, scale, None,
                                                         ctypes.byref(w), ctypes.byref(h)))
        out = np.zeros((h.value, w.value), dtype=np.float32)
        self._check(self._lib.VshipAMD_SSIMU2DebugPlane(self.handle, image, plane, scale,
                                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                         ctypes.byref(w), ctypes.byref(h)))
        return out

    def profile(self, iters: int = 10):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.VshipAMD_SSIMU2Profile(self.handle, iters, ctypes.byref(a), ctypes.byref(b),
                                                      ctypes.byref(c)))
        return a.value, b.value, c.value

    def work(self):
        b, f = ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.VshipAMD_SSIMU2Work(self.handle, ctypes.byref(b), ctypes.byref(f)))
        return b.value, f.value

    def close(self) -> None:
        if getattr(self, "_open", False):
            self._open = False
            check(self._lib.Vship_SSIMU2Free(self.handle))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compute_device_noalloc(src: Colorspace, dis: Colorspace, dptrs1: Sequence[int], dptrs2: Sequence[int],
                           strides1: Sequence[int], strides2: Sequence[int], gpu_id: int = 0) -> float:
    """VshipAMD_ComputeSSIMU2DeviceNoAlloc: the score of one pair of device-resident planes without a
    handler and without allocating device memory (the library's fixed per-GPU scratch array)."""
    lib = load_library()
    p1, l1 = _plane_args(dptrs1, strides1)
    p2, l2 = _plane_args(dptrs2, strides2)
    score = ctypes.c_double(0.0)
    check(lib.VshipAMD_ComputeSSIMU2DeviceNoAlloc(src, dis, ctypes.byref(score), p1, p2, l1, l2, gpu_id))
    return score.value


STORAGE_FP32, STORAGE_FP16 = 0, 1
ALL_SCALES = 0x1F  # scale_mask bit s = pyramid scale s; the score uses scales 0-4


class SSIMU2Batch:
    """A VshipAMD_SSIMU2Batch context: many device-resident pairs per call."""

    def __init__(self, max_pairs: int, storage: int = STORAGE_FP16, gpu_id: int = 0):
        self._lib = load_library()
        self.handle = Handler()
        check(self._lib.VshipAMD_SSIMU2BatchInit(ctypes.byref(self.handle), max_pairs, storage, gpu_id))
        self._open = True
        self.max_pairs = max_pairs

    @staticmethod
    def _args(src: Sequence[Colorspace], dis: Sequence[Colorspace], planes1, planes2, strides1, strides2):
        n = len(src)
        cs1 = (Colorspace * n)(*src)
        cs2 = (Colorspace * n)(*dis)
        p1 = (ctypes.c_void_p * (3 * n))(*[int(x) for trip in planes1 for x in trip])
        p2 = (ctypes.c_void_p * (3 * n))(*[int(x) for trip in planes2 for x in trip])
        l1 = (ctypes.c_int64 * (3 * n))(*[int(x) for trip in strides1 for x in trip])
        l2 = (ctypes.c_int64 * (3 * n))(*[int(x) for trip in strides2 for x in trip])
        return n, cs1, cs2, p1, p2, l1, l2

    def compute(self, src, dis, planes1, planes2, strides1, strides2) -> np.ndarray:
        """Scores of len(src) pairs; planesK[i] are 3 device pointers, stridesK[i] 3 line sizes."""
        n, cs1, cs2, p1, p2, l1, l2 = self._args(src, dis, planes1, planes2, strides1, strides2)
        out = np.zeros(n, dtype=np.float64)
        check(self._lib.VshipAMD_SSIMU2BatchCompute(self.handle, n, cs1, cs2, p1, p2, l1, l2,
                                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def prepare(self, src, dis, planes1, planes2, strides1, strides2):
        """ctypes arguments of one batched call, built once for repeated calls on the same planes (as a
        C caller holds its descriptor and pointer arrays); see compute_prepared."""
        return self._args(src, dis, planes1, planes2, strides1, strides2)

    def compute_prepared(self, args, out: np.ndarray, scale_mask: int = ALL_SCALES) -> np.ndarray:
        """VshipAMD_SSIMU2BatchCompute (or, with scale_mask != ALL_SCALES, ComputeScales) on arguments from
        prepare; scores into out (float64, >= n)."""
        n, cs1, cs2, p1, p2, l1, l2 = args
        dp = out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if scale_mask == ALL_SCALES:
            check(self._lib.VshipAMD_SSIMU2BatchCompute(self.handle, n, cs1, cs2, p1, p2, l1, l2, dp))
        else:
            check(self._lib.VshipAMD_SSIMU2BatchComputeScales(self.handle, n, cs1, cs2, p1, p2, l1, l2, scale_mask,
                                                               dp))
        return out

    def complete_scales(self, pairs, scale_mask: int) -> np.ndarray:
        """VshipAMD_SSIMU2BatchCompleteScales: full scores of pairs (indices into the last compute) after
        blurring the scales in scale_mask on the planes that compute stored."""
        idx = np.ascontiguousarray(pairs, dtype=np.int32)
        out = np.zeros(len(idx), dtype=np.float64)
        check(self._lib.VshipAMD_SSIMU2BatchCompleteScales(self.handle, len(idx),
                                                            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                            scale_mask,
                                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def autotune(self, src, dis, planes1, planes2, strides1, strides2, reps: int = 2):
        n, cs1, cs2, p1, p2, l1, l2 = self._args(src, dis, planes1, planes2, strides1, strides2)
        t = BatchTuning()
        ms = ctypes.c_double()
        check(self._lib.VshipAMD_SSIMU2BatchAutotune(self.handle, n, cs1, cs2, p1, p2, l1, l2, reps,
                                                      ctypes.byref(t), ctypes.byref(ms)))
        return t.as_dict(), ms.value

    def set_tuning(self, blur_config: int, blur_waves: int, front_waves: int, generic_front: int = 0) -> None:
        check(self._lib.VshipAMD_SSIMU2BatchSetTuning(
            self.handle, BatchTuning(blur_config, blur_waves, front_waves, generic_front)))

    def tuning(self) -> dict:
        t = BatchTuning()
        check(self._lib.VshipAMD_SSIMU2BatchGetTuning(self.handle, ctypes.byref(t)))
        return t.as_dict()

    def features(self, npairs: int) -> np.ndarray:
        out = np.zeros((npairs, 108), dtype=np.float64)
        check(self._lib.VshipAMD_SSIMU2BatchFeatures(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def profile(self, iters: int = 5):
        v = [ctypes.c_double() for _ in range(4)]
        check(self._lib.VshipAMD_SSIMU2BatchProfile(self.handle, iters, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def work(self):
        b, f = ctypes.c_double(), ctypes.c_double()
        check(self._lib.VshipAMD_SSIMU2BatchWork(self.handle, ctypes.byref(b), ctypes.byref(f)))
        return b.value, f.value

    def close(self) -> None:
        if getattr(self, "_open", False):
            self._open = False
            check(self._lib.VshipAMD_SSIMU2BatchFree(self.handle))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
