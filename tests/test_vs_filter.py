"""The VapourSynth plugin entry (VapourSynthPluginInit2, vship.SSIMULACRA2,
vship.GpuInfo) driven by the mock VapourSynth host in tests/vs_host/.

CPU tests: registration strings, the RGBS conversion request made to the
host's resize.Bicubic (matrix_in BT.709 above 650 rows, BT.601 below;
reference src/util/torgbs.hpp:8-36) and the argument errors raised before any
GPU work. GPU tests: scores through the filter equal the C API's on the same
float RGB frames, the output frame carries _SSIMULACRA2, and no frame or node
reference leaks. The zimg conversion a real host performs for YUV clips is
not reproduced by the mock (parity unpinned for that step).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from vship_amd import ssimu2 as S
from vship_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOCK = os.path.join(ROOT, "tests", "vs_host", "libmock_vs.so")
CF_GRAY, CF_RGB, CF_YUV = 1, 2, 3
ST_INT, ST_FLOAT = 0, 1


class Host:
    def __init__(self):
        L = ctypes.CDLL(MOCK)
        vp, cp, i64 = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64
        L.mock_load.argtypes = [cp]
        L.mock_plugin_field.restype = cp
        L.mock_function_field.restype = cp
        L.mock_function_field.argtypes = [ctypes.c_int, ctypes.c_int]
        L.mock_clip.restype = vp
        L.mock_clip.argtypes = [ctypes.c_int] * 8
        L.mock_fill.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, i64]
        L.mock_stride.restype = i64
        L.mock_stride.argtypes = [vp, ctypes.c_int]
        L.mock_free_node.argtypes = [vp]
        L.mock_register_resize.argtypes = [vp, vp]
        L.mock_last_resize_args.restype = cp
        L.mock_ssimu2.restype = vp
        L.mock_ssimu2.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, cp, ctypes.c_int]
        L.mock_frame_score.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), cp, ctypes.c_int]
        L.mock_node_shape.argtypes = [vp]
        L.mock_gpu_info.argtypes = [ctypes.c_int, cp, ctypes.c_int]
        self.L = L
        n = L.mock_load(os.path.join(ROOT, "vship_amd", "libvship_amd.so").encode())
        assert n > 0, n

    def rgbs_clip(self, frames):
        """frames: list of (3, h, w) float32 arrays."""
        h, w = frames[0].shape[1:]
        node = self.L.mock_clip(w, h, len(frames), CF_RGB, ST_FLOAT, 32, 0, 0)
        for i, f in enumerate(frames):
            for p in range(3):
                a = np.ascontiguousarray(f[p], dtype=np.float32)
                assert self.L.mock_fill(node, i, p, a.ctypes.data, a.strides[0]) == 0
        return node

    def ssimu2(self, ref, dis, num_stream=-1, gpu_id=-1):
        err = ctypes.create_string_buffer(1024)
        node = self.L.mock_ssimu2(ref, dis, num_stream, gpu_id, err, 1024)
        return node, err.value.decode()

    def score(self, node, n):
        v = ctypes.c_double()
        err = ctypes.create_string_buffer(1024)
        rc = self.L.mock_frame_score(node, n, ctypes.byref(v), err, 1024)
        assert rc == 0, (rc, err.value)
        return v.value


@pytest.fixture(scope="module")
def host():
    return Host()


def test_plugin_registration(host):
    L = host.L
    assert L.mock_plugin_field(0) == b"com.lumen.vship"
    assert L.mock_plugin_field(1) == b"vship"
    assert int(L.mock_plugin_field(3)) == (4 << 16) | 1
    funcs = {L.mock_function_field(i, 0): (L.mock_function_field(i, 1), L.mock_function_field(i, 2))
             for i in range(2)}
    assert funcs[b"SSIMULACRA2"] == (b"reference:vnode;distorted:vnode;numStream:int:opt;gpu_id:int:opt;",
                                     b"clip:vnode;")
    assert funcs[b"GpuInfo"] == (b"gpu_id:int:opt;", b"gpu_human_data:data;")


@pytest.mark.parametrize("h,matrix", [(720, 1), (360, 6)])
def test_yuv_clip_is_sent_to_resize_as_the_reference_does(host, h, matrix):
    L = host.L
    w = h * 16 // 9
    yuv = L.mock_clip(w, h, 1, CF_YUV, ST_INT, 8, 1, 1)
    other = L.mock_clip(w + 2, h, 1, CF_RGB, ST_FLOAT, 32, 0, 0)  # registered conversion result
    L.mock_register_resize(yuv, other)
    rgbs = L.mock_clip(w, h, 1, CF_RGB, ST_FLOAT, 32, 0, 0)
    node, err = host.ssimu2(yuv, rgbs)
    # the converted clip has another size: DifferingInputType, raised before any GPU call
    assert node is None and "DifferingInputType" in err, err
    args = L.mock_last_resize_args().decode()
    assert args == (f"matrix_in={matrix};transfer_in=1;primaries_in=1;range_in=0;matrix=0;transfer=1;"
                    f"primaries=1;range=1;format={(2 << 28) | (1 << 24) | (32 << 16)};")
    L.mock_clear_resize()
    for n in (yuv, other, rgbs):
        L.mock_free_node(n)


def test_argument_errors_before_gpu_work(host):
    L = host.L
    a = L.mock_clip(64, 48, 1, CF_RGB, ST_FLOAT, 32, 0, 0)
    b = L.mock_clip(64, 40, 1, CF_RGB, ST_FLOAT, 32, 0, 0)
    node, err = host.ssimu2(a, b)
    assert node is None and "DifferingInputType" in err
    # a clip the host cannot convert (no resize result registered)
    g = L.mock_clip(64, 48, 1, CF_GRAY, ST_INT, 8, 0, 0)
    node, err = host.ssimu2(g, g)
    assert node is None and "NonRGBSInput" in err
    frames0, nodes0 = L.mock_live_frames(), L.mock_live_nodes()
    for n in (a, b, g):
        L.mock_free_node(n)
    assert L.mock_live_nodes() == nodes0 - 3
    assert L.mock_live_frames() == frames0 - 3


def _float_frames(w, h, n, seed):
    """n (ref, dist) pairs of float RGB frames in [0, 1] from the sRGB
    generator, taken as BT.709-coded values."""
    refs, dists = [], []
    for i in range(n):
        r, d = synth.srgb_pair(w, h, seed=seed + i)
        refs.append(np.stack(r).astype(np.float32) / 255.0)
        dists.append(np.stack(d).astype(np.float32) / 255.0)
    return refs, dists


def _rgbs_colorspace(w, h):
    return S.colorspace(w, h, sample=S.SAMPLE_FLOAT, rng=S.RANGE_FULL, subw=0, subh=0, family=S.FAMILY_RGB,
                        matrix=S.MATRIX_RGB, trc=S.TRC_BT709, prim=S.PRIM_BT709)


@pytest.mark.gpu
def test_filter_scores_equal_the_c_api(host):
    L = host.L
    w, h, n = 200, 120, 3
    refs, dists = _float_frames(w, h, n, seed=40)
    frames0, nodes0 = L.mock_live_frames(), L.mock_live_nodes()
    a, b = host.rgbs_clip(refs), host.rgbs_clip(dists)
    node, err = host.ssimu2(a, b, num_stream=2, gpu_id=0)
    assert node is not None, err
    assert L.mock_node_shape(node) == 0 * 100 + 2 * 10 + 2  # fmParallel, rpStrictSpatial x 2
    got = [host.score(node, i) for i in range(n)]
    cs = _rgbs_colorspace(w, h)
    with S.SSIMU2Handler(cs, cs) as hd:
        want = [hd.compute(list(refs[i]), list(dists[i]), [w * 4] * 3, [w * 4] * 3) for i in range(n)]
    print("filter", got, "C API", want, flush=True)
    assert got == want
    o, _ = oracle.ssimu2(cs, list(refs[0]), [w * 4] * 3, cs, list(dists[0]), [w * 4] * 3)
    assert abs(got[0] - o) <= 0.005
    # identity through the filter
    ident, err = host.ssimu2(a, a)
    assert host.score(ident, 1) == 100.0
    for x in (node, ident, a, b):
        L.mock_free_node(x)
    assert (L.mock_live_frames(), L.mock_live_nodes()) == (frames0, nodes0)


@pytest.mark.gpu
def test_gpu_info(host):
    buf = ctypes.create_string_buffer(4096)
    assert host.L.mock_gpu_info(0, buf, 4096) == 0
    text = buf.value.decode()
    assert "Name:" in text and "PassKernelCheck : 1" in text, text
    assert host.L.mock_gpu_info(-1, buf, 4096) == 0 and buf.value.startswith(b"GPU 0: ")
    assert host.L.mock_gpu_info(99, buf, 4096) == -2 and b"BadDeviceArgument" in buf.value
