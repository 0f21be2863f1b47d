"""GPU test of VshipAMD_ComputeSSIMU2DeviceNoAlloc, the entry point that
allocates no device memory (tables, XYB planes and partial sums in the
library's fixed per-GPU scratch array). Its score must equal the handler
path's (VshipAMD_ComputeSSIMU2Device, same kernels and host completion) bit
for bit, on consecutive calls of different sizes (the scratch keeps the last
call's contents), and frames beyond the array's capacity are rejected on the
host before any device work."""
import numpy as np
import pytest

from vship_amd import ssimu2 as S
from tests import cases

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


class Uploaded:
    def __init__(self, cs, seed, pad=3):
        ref, dist, self.strides = cases.perturbed_pair(cs, seed=seed, pad=pad)
        self.cs = cs
        self.bufs, self.ptrs = [], []
        for img in (ref, dist):
            trip = []
            for p in img:
                b = S.DeviceBuffer(p.nbytes)
                b.upload(p)
                self.bufs.append(b)
                trip.append(b.address)
            self.ptrs.append(trip)

    def handler_score(self):
        with S.SSIMU2Handler(self.cs, self.cs) as h:
            return h.compute_device(self.ptrs[0], self.ptrs[1], self.strides, self.strides)

    def noalloc_score(self):
        return S.compute_device_noalloc(self.cs, self.cs, self.ptrs[0], self.ptrs[1], self.strides, self.strides)

    def free(self):
        for b in self.bufs:
            b.free()


def test_noalloc_equals_handler_path_on_consecutive_calls():
    big = Uploaded(S.colorspace(1920, 1080), seed=11)
    small = Uploaded(S.colorspace(640, 360, sample=S.SAMPLE_UINT10), seed=12)
    try:
        for name, pair in (("1920x1080 8-bit", big), ("640x360 10-bit", small), ("1920x1080 8-bit again", big)):
            _say("no-alloc:", name)
            want = pair.handler_score()
            got = pair.noalloc_score()
            _say(f"  handler {want!r} no-alloc {got!r}")
            assert got == want
        # identical inputs score exactly 100 through the scratch path too
        assert S.compute_device_noalloc(big.cs, big.cs, big.ptrs[0], big.ptrs[0], big.strides, big.strides) == 100.0
    finally:
        big.free()
        small.free()


def test_noalloc_rejects_what_the_scratch_cannot_hold():
    dummy = [1 << 20] * 3  # never dereferenced: the checks come first
    cs = S.colorspace(2560, 1440)
    with pytest.raises(S.VshipError) as e:
        S.compute_device_noalloc(cs, cs, dummy, dummy, [2560, 1280, 1280], [2560, 1280, 1280])
    assert e.value.code == 1 and "at most" in str(e.value)
    cs = S.colorspace(640, 360)
    cs.crop.left = 8
    with pytest.raises(S.VshipError) as e:
        S.compute_device_noalloc(cs, cs, dummy, dummy, [640, 320, 320], [640, 320, 320])
    assert e.value.code == 5
