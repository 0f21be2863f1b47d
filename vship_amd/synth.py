"""Seeded synthetic frame pairs for tests and benchmarks (numpy PCG64).

Content generators follow the workload descriptions used to measure this
path:

* ``srgb_pair`` - 8-bit planar RGB: 128 + a sum of 32 random 2-D sinusoids
  (|f| ~ U[0.005, 0.2] cycles/px, orientation and phase ~ U[0, 2pi),
  amplitude ~ U[5, 40]) + N(0, 4^2) noise, drawn independently per channel;
  distorted = Gaussian blur sigma 1.0 (scipy.ndimage.gaussian_filter, mode
  'reflect', truncate 4) + N(0, 3^2) noise, rounded and clipped.
* ``yuv_sequence`` - YUV 4:2:0 frames (8-bit or 10-bit limited range):
  luma = sinusoid sum mapped into [16, 235] (x4 for 10-bit), chroma = smooth
  random fields (Gaussian-filtered white noise, sigma 24 chroma px) around 128
  within [16, 240], the whole content translated 2 luma px per frame.
  Distorted = every plane 8x8 block DCT-II (orthonormal) quantised with one
  step q per frame (default q ~ U{4..32}) and reconstructed; then the chroma
  planes are 2x box-downsampled and 2x upsampled again (each output sample =
  its 2x2 parent; phase-aligned replication) before rounding and clipping.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def _sinusoid_sum(h: int, w: int, rng: np.random.Generator, n: int = 32, fmin: float = 0.005,
                  fmax: float = 0.2, amin: float = 5.0, amax: float = 40.0) -> np.ndarray:
    """Sum of n sinusoids on an h x w grid as one rank-2n matrix product."""
    f = rng.uniform(fmin, fmax, n)
    theta = rng.uniform(0, 2 * np.pi, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    amp = rng.uniform(amin, amax, n)
    ax = 2 * np.pi * f * np.cos(theta)
    by = 2 * np.pi * f * np.sin(theta)
    x = np.arange(w, dtype=np.float64)
    y = np.arange(h, dtype=np.float64)
    # sin(ax + by + phi) = sin(ax + phi) cos(by) + cos(ax + phi) sin(by)
    X = np.concatenate([np.sin(np.outer(ax, x) + phi[:, None]), np.cos(np.outer(ax, x) + phi[:, None])], 0)
    Y = np.concatenate([np.cos(np.outer(y, by)) * amp, np.sin(np.outer(y, by)) * amp], 1)
    return Y @ X


def srgb_pair(width: int = 256, height: int = 256, seed: int = 0):
    """(ref, dist): lists of 3 uint8 planes of shape (height, width)."""
    rng = np.random.default_rng(seed)
    ref, dist = [], []
    for _ in range(3):
        base = 128.0 + _sinusoid_sum(height, width, rng) + rng.normal(0, 4.0, (height, width))
        r = np.clip(np.round(base), 0, 255)
        d = ndimage.gaussian_filter(r, 1.0, mode="reflect", truncate=4.0) + rng.normal(0, 3.0, (height, width))
        ref.append(r.astype(np.uint8))
        dist.append(np.clip(np.round(d), 0, 255).astype(np.uint8))
    return ref, dist


_DCT8 = np.array([[np.sqrt((1 if k == 0 else 2) / 8.0) * np.cos((2 * n + 1) * k * np.pi / 16) for n in range(8)]
                  for k in range(8)])


def dct_quantize(plane: np.ndarray, q: float) -> np.ndarray:
    """8x8 orthonormal DCT-II, uniform quantisation with step q, inverse DCT."""
    h, w = plane.shape
    H, W = (h + 7) // 8 * 8, (w + 7) // 8 * 8
    p = np.pad(plane.astype(np.float64), ((0, H - h), (0, W - w)), mode="edge")
    blocks = p.reshape(H // 8, 8, W // 8, 8).transpose(0, 2, 1, 3)
    coef = _DCT8 @ blocks @ _DCT8.T
    coef = np.round(coef / q) * q
    rec = _DCT8.T @ coef @ _DCT8
    return rec.transpose(0, 2, 1, 3).reshape(H, W)[:h, :w]


def _resample_chroma(c: np.ndarray) -> np.ndarray:
    h, w = c.shape
    H, W = (h + 1) // 2 * 2, (w + 1) // 2 * 2
    p = np.pad(c, ((0, H - h), (0, W - w)), mode="edge")
    small = p.reshape(H // 2, 2, W // 2, 2).mean(axis=(1, 3))
    return np.repeat(np.repeat(small, 2, 0), 2, 1)[:h, :w]


class YuvSequence:
    """Translating synthetic YUV 4:2:0 content and its DCT-quantised version."""

    def __init__(self, width: int, height: int, frames: int, bits: int = 8, seed: int = 0, shift: int = 2):
        self.w, self.h, self.n, self.bits, self.shift = width, height, frames, bits, shift
        rng = np.random.default_rng(seed)
        self.rng = rng
        span = width + shift * max(frames - 1, 0)
        luma = _sinusoid_sum(height, span, rng)
        lo, hi = luma.min(), luma.max()
        self.luma = 16.0 + (luma - lo) * (219.0 / max(hi - lo, 1e-9))
        cw, ch = (span + 1) // 2, (height + 1) // 2
        self.chroma = []
        for _ in range(2):
            f = ndimage.gaussian_filter(rng.normal(0, 1, (ch, cw)), 24.0, mode="wrap")
            f = f / max(np.abs(f).max(), 1e-9)
            self.chroma.append(np.clip(128.0 + 100.0 * f, 16.0, 240.0))

    def _quant(self, v: np.ndarray) -> np.ndarray:
        scale = 1 << (self.bits - 8)
        top = (1 << self.bits) - 1
        dt = np.uint8 if self.bits == 8 else np.uint16
        return np.clip(np.round(v * scale), 0, top).astype(dt)

    def frame(self, i: int, q: float | None = None):
        """(ref planes, dist planes, q) of frame i."""
        x0 = i * self.shift
        y = self.luma[:, x0:x0 + self.w]
        cw, ch = (self.w + 1) // 2, (self.h + 1) // 2
        cx0 = x0 // 2
        u = self.chroma[0][:ch, cx0:cx0 + cw]
        v = self.chroma[1][:ch, cx0:cx0 + cw]
        ref = [self._quant(y), self._quant(u), self._quant(v)]
        if q is None:
            q = float(np.random.default_rng([i, 12345]).integers(4, 33))
        scale = 1 << (self.bits - 8)
        dy = dct_quantize(ref[0].astype(np.float64) / scale, q)
        du = _resample_chroma(dct_quantize(ref[1].astype(np.float64) / scale, q))
        dv = _resample_chroma(dct_quantize(ref[2].astype(np.float64) / scale, q))
        dist = [self._quant(dy), self._quant(du), self._quant(dv)]
        return ref, dist, q


def random_planes(cs, rng: np.random.Generator, pad: int = 0):
    """Random planes matching a Colorspace descriptor (optionally row-padded).

    Returns (planes, strides): each plane is a 2-D uint8 view of shape
    (rows, stride) whose first width*bytes bytes per row hold samples.
    """
    from .ssimu2 import SAMPLE_FLOAT, SAMPLE_HALF, SAMPLE_UINT8
    planes, strides = [], []
    nbytes = cs.sample_bytes
    bits = {3: 9, 5: 10, 7: 12, 9: 14, 11: 16}.get(cs.sample, 8)
    for p in range(3):
        h, w = cs.plane_shape(p)
        if cs.sample == SAMPLE_UINT8:
            vals = rng.integers(0, 256, (h, w)).astype(np.uint8)
        elif cs.sample == SAMPLE_FLOAT:
            vals = rng.uniform(0, 1, (h, w)).astype(np.float32)
        elif cs.sample == SAMPLE_HALF:
            vals = rng.uniform(0, 1, (h, w)).astype(np.float16)
        else:
            vals = rng.integers(0, 1 << bits, (h, w)).astype(np.uint16)
        raw = vals.view(np.uint8).reshape(h, w * nbytes)
        stride = w * nbytes + pad * (p + 1)
        buf = np.zeros((h, stride), dtype=np.uint8)
        buf[:, :w * nbytes] = raw
        planes.append(buf)
        strides.append(stride)
    return planes, strides


# Mixed-resolution batch: 3840x2160 (10-bit), 1920x1080, 1280x720 and 640x360
# (8-bit) YUV 4:2:0 BT.709 limited pairs in the ratio 1:4:8:16.
MIXED_SIZES = ((3840, 2160, 10), (1920, 1080, 8), (1280, 720, 8), (640, 360, 8))
MIXED_RATIO = (1, 4, 8, 16)


def mixed_counts(total: int):
    """Pairs per size for `total` pairs split 1:4:8:16 by largest remainders
    (512 -> 18, 71, 141, 282)."""
    s = sum(MIXED_RATIO)
    exact = [total * r / s for r in MIXED_RATIO]
    counts = [int(e) for e in exact]
    order = sorted(range(len(exact)), key=lambda i: exact[i] - counts[i], reverse=True)
    for i in order[: total - sum(counts)]:
        counts[i] += 1
    return counts


_SEQS = {}  # sequences of the jobs of frames_in_processes (inherited by forked workers)


def _seq_frame(job):
    key, i = job
    seq = _SEQS[key]
    ref, dist, q = seq.frame(i)
    return seq.w, seq.h, seq.bits, ref, dist, q


def frames_in_processes(seqs: dict, jobs, workers: int = 16):
    """[(w, h, bits, ref, dist, q)] of seqs[key].frame(i) for (key, i) jobs,
    in forked worker processes (numpy holds the GIL for much of a frame).
    Call it before the process uses the GPU: the workers are forks."""
    import concurrent.futures as cf
    import multiprocessing as mp

    _SEQS.clear()
    _SEQS.update(seqs)
    try:
        with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork")) as ex:
            return list(ex.map(_seq_frame, jobs, chunksize=max(1, len(jobs) // (4 * workers))))
    finally:
        _SEQS.clear()


def mixed_batch(counts, seed: int = 0, workers: int = 8, processes: bool = False):
    """Seeded, shuffled list of (width, height, bits, ref planes, dist planes, q)
    with counts[k] pairs of MIXED_SIZES[k]; the content and distortion are those
    of YuvSequence. processes: generate in forked worker processes (see
    frames_in_processes), otherwise in threads."""
    import concurrent.futures as cf

    seqs, jobs = {}, []
    for k, ((w, h, bits), n) in enumerate(zip(MIXED_SIZES, counts)):
        if n:
            seqs[k] = YuvSequence(w, h, n, bits=bits, seed=seed + 101 * k)
            jobs += [(k, i) for i in range(n)]
    if processes:
        out = frames_in_processes(seqs, jobs, workers)
    else:
        def one(job):
            k, i = job
            ref, dist, q = seqs[k].frame(i)
            return (seqs[k].w, seqs[k].h, seqs[k].bits, ref, dist, q)

        with cf.ThreadPoolExecutor(max_workers=workers) as ex:
            out = list(ex.map(one, jobs))
    order = np.random.default_rng(seed).permutation(len(out))
    return [out[i] for i in order]
