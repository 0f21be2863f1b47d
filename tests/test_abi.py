"""CPU checks of the C ABI: the library loads, exports every entry point the
headers declare, keeps the reference's struct layouts, and honours the error
and message contracts (reference src/VshipLib.cpp:102-286) without a GPU."""
import ctypes
import os
import re

import pytest

from vship_amd import ssimu2 as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAS_GPU = os.path.exists("/dev/kfd")


def declared_functions():
    names = []
    for hdr in ("VshipAPI.h", "vship_amd.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        names += re.findall(r"EXPORTPREPROCESS\s+[\w\s\*]+?\b(\w+)\s*\(", text)
    return names


def test_library_exports_every_declared_entry_point():
    lib = S.load_library()
    names = declared_functions()
    assert len(names) >= 29 + 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_symbol_set_is_covered():
    # the 29 Vship_* entry points of the reference header (src/VshipAPI.h:35-190)
    ref = ["GetVersion", "GetDeviceCount", "SetDevice", "GetDeviceInfo", "GPUFullCheck", "GetErrorMessage",
           "GetDetailedLastError", "PinnedMalloc", "PinnedFree", "SSIMU2Init", "SSIMU2Init2", "SSIMU2Free",
           "ComputeSSIMU2", "SSIMU2GetDetailedLastError", "ButteraugliInit", "ButteraugliInit2",
           "ButteraugliFree", "ComputeButteraugli", "ButteraugliGetDetailedLastError", "CVVDPInit", "CVVDPInit2",
           "CVVDPInit3", "CVVDPFree", "CVVDPGetDistmapResolution", "ResetCVVDP", "ResetScoreCVVDP",
           "LoadTemporalCVVDP", "ComputeCVVDP", "CVVDPGetDetailedLastError"]
    lib = S.load_library()
    assert all(hasattr(lib, "Vship_" + n) for n in ref)


def test_struct_layouts():
    assert ctypes.sizeof(S.Colorspace) == 88
    assert S.Colorspace.sample.offset == 32 and S.Colorspace.crop.offset == 68
    assert ctypes.sizeof(S.DeviceInfo) == 280
    assert ctypes.sizeof(S.Handler) == 4


def test_version():
    v = S.load_library().Vship_GetVersion()
    assert (v.major, v.minor, v.minorMinor, v.backend) == (4, 1, 0, 0)


def test_error_message_size_contract():
    lib = S.load_library()
    for code in S.ERROR_NAMES:
        n = lib.Vship_GetErrorMessage(code, None, 0)
        full = S.error_message(code)
        assert n == len(full) + 1
        assert full.startswith(("InternalError" if code == 12 else S.ERROR_NAMES[code]) + ":")
        buf = ctypes.create_string_buffer(8)
        assert lib.Vship_GetErrorMessage(code, buf, 8) == n
        assert buf.value.decode() == full[:7]


def test_bad_handler_paths():
    lib = S.load_library()
    bad = S.Handler(12345)
    assert lib.Vship_SSIMU2Free(bad) == 10
    n = lib.Vship_GetDetailedLastError(None, 0)
    buf = ctypes.create_string_buffer(n)
    lib.Vship_GetDetailedLastError(buf, n)
    msg = buf.value.decode()
    assert msg.startswith("VshipException\nBadHandler:") and " - At line " in msg
    assert lib.Vship_SSIMU2GetDetailedLastError(bad, None, 0) == 0
    score = ctypes.c_double()
    planes = (ctypes.c_void_p * 3)()
    ls = (ctypes.c_int64 * 3)()
    assert lib.Vship_ComputeSSIMU2(bad, ctypes.byref(score), planes, planes, ls, ls) == 10


def test_other_metrics_are_declared_but_absent():
    lib = S.load_library()
    h = ctypes.c_uint(0)
    lib.Vship_ButteraugliInit2.restype = ctypes.c_int
    cs = S.colorspace(16, 16)
    lib.Vship_ButteraugliInit2.argtypes = [ctypes.POINTER(ctypes.c_uint), S.Colorspace, S.Colorspace, ctypes.c_int,
                                           ctypes.c_float, ctypes.c_int]
    assert lib.Vship_ButteraugliInit2(ctypes.byref(h), cs, cs, 5, 203.0, 0) == 15


@pytest.mark.skipif(HAS_GPU, reason="checks the behaviour without a GPU")
def test_no_gpu_errors():
    lib = S.load_library()
    n = ctypes.c_int(-1)
    rc = lib.Vship_GetDeviceCount(ctypes.byref(n))
    assert rc in (6, 7)  # DeviceCountError / NoDeviceDetected
    assert lib.Vship_GPUFullCheck(0) in (6, 7)
    with pytest.raises(S.VshipError) as e:
        S.SSIMU2Handler(S.colorspace(64, 64), S.colorspace(64, 64))
    assert e.value.code in (6, 7)


def test_colourspace_validation_happens_before_device_use():
    # invalid enum values are rejected by the host-side validation (NonRGBSInput)
    # or, without a GPU, by the device check first; never a crash
    cs = S.colorspace(64, 64)
    cs.sample = 4  # not a Vship_Sample_t value
    with pytest.raises(S.VshipError) as e:
        S.SSIMU2Handler(cs, cs)
    assert e.value.code in (5, 6, 7)


def test_pinned_free_of_garbage_pointer_is_an_error_not_a_crash():
    if HAS_GPU:
        pytest.skip("only meaningful without a GPU runtime")
    lib = S.load_library()
    assert lib.Vship_PinnedFree(ctypes.c_void_p(0x1234)) in (11,)


def test_free_from_a_second_thread_while_computing():
    # Without a GPU the init fails but hands out an id (as the reference does),
    # so the handle table is exercised through the real entry points: one
    # thread computes on ids that another frees and re-creates. Every call
    # must return an error code (BadHandler, or the init's device error),
    # never crash. On a GPU box the handlers are real and compute fails on
    # the null planes (BadPointer) instead.
    import threading
    lib = S.load_library()
    cs = S.colorspace(64, 64)
    ids = []
    for _ in range(4):
        h = S.Handler()
        lib.Vship_SSIMU2Init2(ctypes.byref(h), cs, cs, 0)
        ids.append(h.id)
    stop = threading.Event()
    codes = set()

    def compute():
        score = ctypes.c_double()
        planes = (ctypes.c_void_p * 3)()
        ls = (ctypes.c_int64 * 3)(64, 32, 32)
        k = 0
        while not stop.is_set():
            codes.add(lib.Vship_ComputeSSIMU2(S.Handler(ids[k % 4]), ctypes.byref(score), planes, planes, ls, ls))
            k += 1

    def churn():
        for it in range(2000):
            j = it % 4
            lib.Vship_SSIMU2Free(S.Handler(ids[j]))
            h = S.Handler()
            lib.Vship_SSIMU2Init2(ctypes.byref(h), cs, cs, 0)
            ids[j] = h.id

    workers = [threading.Thread(target=compute) for _ in range(3)]
    for w in workers:
        w.start()
    churn()
    stop.set()
    for w in workers:
        w.join()
    for i in ids:
        lib.Vship_SSIMU2Free(S.Handler(i))
    assert codes and codes <= {6, 7, 10, 11}, codes


def test_noalloc_entry_point_rejects_on_the_host():
    # frames beyond the no-allocation entry point's fixed scratch array, and
    # crop/resize, are refused before any device is touched (also without a GPU)
    dummy = [1 << 20] * 3
    cs = S.colorspace(2560, 1440)
    with pytest.raises(S.VshipError) as e:
        S.compute_device_noalloc(cs, cs, dummy, dummy, [2560, 1280, 1280], [2560, 1280, 1280])
    assert e.value.code == 1 and "at most 33554432 bytes" in str(e.value)
    cs = S.colorspace(640, 360)
    cs.target_width = 320
    with pytest.raises(S.VshipError) as e:
        S.compute_device_noalloc(cs, cs, dummy, dummy, [640, 320, 320], [640, 320, 320])
    assert e.value.code == 5
