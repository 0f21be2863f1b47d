"""ctypes access to the reference libvship built from its own sources into
oracle/_ref/ by oracle/build_ref.sh (a checker, used only by tests)."""
from __future__ import annotations

import ctypes
import os

from vship_amd.ssimu2 import Colorspace, Handler, _plane_args

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libvship_ref.so")


def available() -> bool:
    return os.path.exists(REF_LIB)


class RefVship:
    def __init__(self, path: str = REF_LIB):
        self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        P = ctypes.POINTER
        L = self.lib
        L.Vship_SSIMU2Init2.restype = ctypes.c_int
        L.Vship_SSIMU2Init2.argtypes = [P(Handler), Colorspace, Colorspace, ctypes.c_int]
        L.Vship_SSIMU2Free.restype = ctypes.c_int
        L.Vship_SSIMU2Free.argtypes = [Handler]
        L.Vship_ComputeSSIMU2.restype = ctypes.c_int
        L.Vship_ComputeSSIMU2.argtypes = [Handler, P(ctypes.c_double), P(ctypes.c_void_p), P(ctypes.c_void_p),
                                          P(ctypes.c_int64), P(ctypes.c_int64)]
        L.Vship_PinnedMalloc.restype = ctypes.c_int
        L.Vship_PinnedMalloc.argtypes = [P(ctypes.c_void_p), ctypes.c_uint64]

    def handler(self, src: Colorspace, dis: Colorspace, gpu: int = 0) -> Handler:
        h = Handler()
        rc = self.lib.Vship_SSIMU2Init2(ctypes.byref(h), src, dis, gpu)
        if rc != 0:
            raise RuntimeError(f"reference Vship_SSIMU2Init2 failed: {rc}")
        return h

    def compute(self, h: Handler, planes1, strides1, planes2, strides2) -> float:
        p1, l1 = _plane_args(planes1, strides1)
        p2, l2 = _plane_args(planes2, strides2)
        s = ctypes.c_double()
        rc = self.lib.Vship_ComputeSSIMU2(h, ctypes.byref(s), p1, p2, l1, l2)
        if rc != 0:
            raise RuntimeError(f"reference Vship_ComputeSSIMU2 failed: {rc}")
        return s.value

    def free(self, h: Handler) -> None:
        self.lib.Vship_SSIMU2Free(h)

    def score(self, src: Colorspace, dis: Colorspace, planes1, strides1, planes2, strides2) -> float:
        h = self.handler(src, dis)
        try:
            return self.compute(h, planes1, strides1, planes2, strides2)
        finally:
            self.free(h)
