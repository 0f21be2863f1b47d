"""The throughput driver vship_amd/ffvship_amd (FFVship pipeline over raw YUV
files). CPU: frame selection against the examples of the reference's
doc/FFVship.md ("encoded-offset", "source-indices"), argument errors. GPU:
per-frame scores in --json and --live-score-output equal the C API's on the
same frames and are within the plain path's 0.005 of the FP64 oracle
(8-bit, 10-bit, cropped)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from vship_amd import ssimu2 as S
from vship_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "vship_amd", "ffvship_amd")


def run(*args, check=True):
    p = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=240)
    if check and p.returncode != 0:
        raise AssertionError(f"driver failed ({p.returncode}): {p.stderr}")
    return p


def write_raw(path, frames):
    with open(path, "wb") as f:
        for planes in frames:
            for p in planes:
                f.write(np.ascontiguousarray(p).tobytes())


def tiny_files(tmp_path, n_src, n_enc):
    frame = [np.zeros((16, 16), np.uint8), np.zeros((8, 8), np.uint8), np.zeros((8, 8), np.uint8)]
    a, b = tmp_path / "src.yuv", tmp_path / "enc.yuv"
    write_raw(a, [frame] * n_src)
    write_raw(b, [frame] * n_enc)
    return a, b


def pairs(out):
    lines = out.strip().splitlines()
    n = int(lines[0])
    got = [tuple(map(int, l.split())) for l in lines[1:]]
    assert len(got) == n
    return got


def test_frame_selection_follows_the_reference_doc(tmp_path):
    a, b = tiny_files(tmp_path, 5, 5)
    common = ["--source", a, "--encoded", b, "--width", 16, "--height", 16, "--dry-run"]
    assert pairs(run(*common, "--start", 0, "--end", 5, "--encoded-offset", 1).stdout) == \
        [(0, 1), (1, 2), (2, 3), (3, 4)]
    assert pairs(run(*common, "--start", 0, "--end", 5, "--encoded-offset", -1).stdout) == \
        [(1, 0), (2, 1), (3, 2), (4, 3)]
    assert pairs(run(*common, "--start", 1, "--every", 2).stdout) == [(1, 1), (3, 3)]
    a, b = tiny_files(tmp_path, 40, 40)
    common = ["--source", a, "--encoded", b, "--width", 16, "--height", 16, "--dry-run"]
    # "--start 10 --every 2 --encoded-offset 1 --source-indices 13: frame 36 and 37"
    assert pairs(run(*common, "--start", 10, "--every", 2, "--encoded-offset", 1, "--source-indices", 13).stdout) \
        == [(36, 37)]
    assert pairs(run(*common, "--source-indices", "1,2", "--encoded-indices", "3,4").stdout) == [(1, 3), (2, 4)]


def test_argument_errors(tmp_path):
    a, b = tiny_files(tmp_path, 2, 2)
    assert run("--help").returncode == 0
    assert run("--source", a, check=False).returncode == 2
    assert run("--source", a, "--encoded", b, "--width", 16, "--height", 16, "--pix-fmt", "nv12",
               check=False).returncode == 2
    p = run("--source", a, "--encoded", b, "--width", 16, "--height", 16, "--encoded-indices", "1", "--dry-run",
            check=False)
    assert p.returncode == 2 and "source-indices" in p.stderr
    assert run("--source", a, "--encoded", b, "--width", 16, "--height", 16, "--bogus", check=False).returncode == 2


def _frames(w, h, n, bits):
    seq = synth.YuvSequence(w, h, n, bits=bits, seed=9)
    return [seq.frame(i) for i in range(n)]


@pytest.mark.gpu
@pytest.mark.parametrize("bits,crop", [(8, None), (10, None), (8, (2, 4, 10, 6))])
def test_driver_scores_equal_the_c_api(tmp_path, bits, crop):
    w, h, n = 320, 180, 6
    frames = _frames(w, h, n, bits)
    src, enc = tmp_path / "s.yuv", tmp_path / "e.yuv"
    write_raw(src, [f[0] for f in frames])
    write_raw(enc, [f[1] for f in frames])
    fmt = "yuv420p" if bits == 8 else "yuv420p10le"
    extra = []
    cs = S.colorspace(w, h, sample=S.SAMPLE_UINT8 if bits == 8 else S.SAMPLE_UINT10)
    if crop:
        t, bo, l, r = crop
        extra = ["--cropTopSource", t, "--cropBottomSource", bo, "--cropLeftSource", l, "--cropRightSource", r,
                 "--cropTopEncoded", t, "--cropBottomEncoded", bo, "--cropLeftEncoded", l, "--cropRightEncoded", r]
        cs.crop = S.CropRectangle(*crop)
    out = tmp_path / "scores.json"
    p = run("--source", src, "--encoded", enc, "--width", w, "--height", h, "--pix-fmt", fmt, "-g", 2, "-t", 2,
            "--json", out, *extra)
    print(p.stdout, flush=True)
    got = [v[0] for v in json.load(open(out))]
    with S.SSIMU2Handler(cs, cs) as hd:
        want = [hd.compute(f[0], f[1], S.host_strides(f[0]), S.host_strides(f[1])) for f in frames]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    exact = [oracle.ssimu2(cs, f[0], S.host_strides(f[0]), cs, f[1], S.host_strides(f[1]))[0] for f in frames]
    print("driver vs oracle", max(abs(a - b) for a, b in zip(got, exact)), flush=True)
    np.testing.assert_allclose(got, exact, rtol=0, atol=0.005)
    assert "Computed 6 frames" in p.stdout and "Average" in p.stdout
    live = run("--source", src, "--encoded", enc, "--width", w, "--height", h, "--pix-fmt", fmt, "-g", 3,
               "--live-score-output", *extra).stdout.strip().splitlines()
    assert int(live[0]) == n and len(live) == n + 1
    got_live = dict((int(l.split()[0]), float(l.split()[1])) for l in live[1:])
    np.testing.assert_allclose([got_live[i] for i in range(n)], want, rtol=1e-13)
