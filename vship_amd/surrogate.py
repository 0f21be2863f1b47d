"""Score surrogate: an MLP that predicts the SSIMULACRA2 score from 54
features (those of three of the five scales) plus log(pixels), with an
uncertainty head, trained on the GPU by the library's own kernels
(VshipAMD_MLP*, vship_amd/csrc/mlp.hip). Two networks ship: one on the
scale 0-2 features (data/surrogate.npz) and one on the scale 2-4 features
(data/surrogate_coarse.npz), whose inputs cost a small part of the blur, so
that screening with it is faster than exact-only screening.

Pieces:
* ``SurrogateNet`` - ctypes owner of one network (BF16 weights, FP32 math).
* ``TrainingPairs`` - synthetic YUV 4:2:0 8-bit BT.709 pairs from 256x256 to
  1920x1080: sinusoid-sum content, then 8x8 DCT quantisation (q log-uniform
  in [2, 64]), Gaussian blur (sigma 0 with probability 1/3, else uniform in
  [0, 2]) and Gaussian noise (sigma 0 with probability 1/3, else uniform in
  [0, 10]) on every plane; half of the pairs also get the chroma planes 2x
  down- and up-sampled after quantisation.
* ``label`` - exact features and scores of pairs from the batched metric.
* ``screen`` - target-quality screening: the network ranks candidates, the
  exact metric confirms the top candidates of each source, and candidates
  whose predicted standard deviation is above a threshold go to the exact
  metric directly (fallback).

``python -m vship_amd.surrogate train ...`` generates data with the metric,
trains, evaluates on held-out pairs and writes the weights and a report.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes
import json
import math
import os
import sys
import time

import numpy as np
from scipy import ndimage

from . import ssimu2 as S
from . import synth

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
PARAMS_FILE = os.path.join(_DATA, "surrogate.npz")  # inputs: scale 0-2 features
COARSE_PARAMS_FILE = os.path.join(_DATA, "surrogate_coarse.npz")  # inputs: scale 2-4 features
N_IN, N_HID, N_OUT = 55, 128, 2
# Scale masks of the networks' inputs. Scales 0-2 hold 98% of the pyramid's
# pixels; scales 2-4 hold 6%, so a screening pass that computes them for
# every candidate costs the front-end and a small part of the blur, and
# confirming a candidate exactly then blurs its scales 0-1 from the stored
# planes.
FEATURE_SCALES = 0b00111
COARSE_SCALES = 0b11100


def feature_index(scales: int = FEATURE_SCALES) -> np.ndarray:
    """Indices plane * 36 + scale * 6 + k into the 108 features of the 54
    inputs of the scales in the mask (three of the five)."""
    idx = [p * 36 + s * 6 + k for p in range(3) for s in range(5) if scales >> s & 1 for k in range(6)]
    if len(idx) != 54:
        raise ValueError("a network's inputs are the features of three scales")
    return np.array(idx)


FEATURE_INDEX = feature_index(FEATURE_SCALES)


def _bind(lib):
    if getattr(lib, "_mlp_bound", False):
        return lib
    P = ctypes.POINTER
    E = ctypes.c_int
    f = P(ctypes.c_float)
    sigs = {
        "VshipAMD_MLPParamCount": (ctypes.c_int, []),
        "VshipAMD_MLPInit": (E, [P(S.Handler), ctypes.c_int, ctypes.c_int]),
        "VshipAMD_MLPFree": (E, [S.Handler]),
        "VshipAMD_MLPSetParams": (E, [S.Handler, f, ctypes.c_int]),
        "VshipAMD_MLPGetParams": (E, [S.Handler, f, ctypes.c_int]),
        "VshipAMD_MLPSetNorm": (E, [S.Handler, f, f, ctypes.c_float, ctypes.c_float]),
        "VshipAMD_MLPPredict": (E, [S.Handler, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
        "VshipAMD_MLPGradients": (E, [S.Handler, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, f,
                                      P(ctypes.c_double)]),
        "VshipAMD_MLPTrain": (E, [S.Handler, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_uint64, f]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib._mlp_bound = True
    return lib


def inputs_from_features(features108, pixels, scales: int = FEATURE_SCALES) -> np.ndarray:
    """(n, 55) float32 network inputs: the features of the scales in the mask
    and log(pixels)."""
    f = np.asarray(features108, dtype=np.float64).reshape(-1, 108)
    x = np.empty((f.shape[0], N_IN), dtype=np.float32)
    x[:, :54] = f[:, feature_index(scales)]
    x[:, 54] = np.log(np.asarray(pixels, dtype=np.float64))
    return x


def init_params(seed: int = 0) -> np.ndarray:
    """Seeded initial parameters: N(0, 1/fan_in) weights, zero biases, and a
    smaller last layer so the first predictions sit near the target mean."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, 1 / math.sqrt(N_IN), (N_HID, N_IN))
    w2 = rng.normal(0, 1 / math.sqrt(N_HID), (N_HID, N_HID))
    w3 = rng.normal(0, 0.1 / math.sqrt(N_HID), (N_OUT, N_HID))
    return np.concatenate([w1.ravel(), np.zeros(N_HID), w2.ravel(), np.zeros(N_HID), w3.ravel(),
                           np.zeros(N_OUT)]).astype(np.float32)


class _Dev:
    """Device copy of a float32 host array."""

    def __init__(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        self.n = a.size
        self.buf = S.DeviceBuffer(max(4, a.nbytes))
        if a.nbytes:
            self.buf.upload(a)

    @property
    def ptr(self):
        return ctypes.c_void_p(self.buf.address)

    def get(self, count=None):
        count = self.n if count is None else count
        return self.buf.download(4 * count).view(np.float32).copy()

    def free(self):
        self.buf.free()


class SurrogateNet:
    """One network on one GPU (VshipAMD_MLP handle)."""

    def __init__(self, max_batch: int = 4096, gpu_id: int = 0):
        self._lib = _bind(S.load_library())
        self.handle = S.Handler()
        S.check(self._lib.VshipAMD_MLPInit(ctypes.byref(self.handle), max_batch, gpu_id))
        self._open = True
        self.max_batch = max_batch
        self.param_count = self._lib.VshipAMD_MLPParamCount()
        self.norm = None
        self.feature_scales = FEATURE_SCALES  # scale mask of the inputs (set by load)

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        S.check(self._lib.VshipAMD_MLPSetParams(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))

    def params(self) -> np.ndarray:
        p = np.empty(self.param_count, dtype=np.float32)
        S.check(self._lib.VshipAMD_MLPGetParams(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))
        return p

    def set_norm(self, xmean, xstd, ymean: float, ystd: float) -> None:
        xm = np.ascontiguousarray(xmean, dtype=np.float32)
        xs = np.ascontiguousarray(xstd, dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        S.check(self._lib.VshipAMD_MLPSetNorm(self.handle, xm.ctypes.data_as(fp), xs.ctypes.data_as(fp),
                                              float(ymean), float(ystd)))
        self.norm = (xm, xs, np.float32(ymean), np.float32(ystd))

    def fit_norm(self, X, y) -> None:
        X = np.asarray(X, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        xs = X.std(0)
        xs[xs < 1e-9] = 0.0
        self.set_norm(X.mean(0), xs, y.mean(), max(y.std(), 1e-6))

    def predict(self, X):
        """(mean, sigma) score predictions of host samples X (n, 55)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        dx, dm, ds = _Dev(X), _Dev(np.zeros(n)), _Dev(np.zeros(n))
        try:
            S.check(self._lib.VshipAMD_MLPPredict(self.handle, dx.ptr, n, dm.ptr, ds.ptr))
            return dm.get(n).astype(np.float64), ds.get(n).astype(np.float64)
        finally:
            for d in (dx, dm, ds):
                d.free()

    def gradients(self, X, Y):
        """(mean loss, gradient) over the samples, no update."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        dx, dy = _Dev(X), _Dev(Y)
        g = np.empty(self.param_count, dtype=np.float32)
        loss = ctypes.c_double()
        try:
            S.check(self._lib.VshipAMD_MLPGradients(self.handle, dx.ptr, dy.ptr, X.shape[0],
                                                    g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                    ctypes.byref(loss)))
        finally:
            dx.free()
            dy.free()
        return loss.value, g

    def train(self, X, Y, epochs: int = 40, batch: int = 1024, lr: float = 2e-3, weight_decay: float = 0.0,
              seed: int = 0) -> np.ndarray:
        """Adam training on the GPU; returns the mean loss of each epoch."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        dx, dy = _Dev(X), _Dev(Y)
        losses = np.empty(epochs, dtype=np.float32)
        try:
            S.check(self._lib.VshipAMD_MLPTrain(self.handle, dx.ptr, dy.ptr, X.shape[0], epochs, batch, lr,
                                                weight_decay, seed,
                                                losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        finally:
            dx.free()
            dy.free()
        return losses

    def save(self, path: str, **meta) -> None:
        xm, xs, ym, ys = self.norm
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        meta = dict(meta, feature_scales=int(self.feature_scales))
        np.savez(path, params=self.params(), xmean=xm, xstd=xs, ymean=ym, ystd=ys,
                 meta=np.array(json.dumps(meta)))

    def load(self, path: str) -> dict:
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])
            self.set_norm(z["xmean"], z["xstd"], float(z["ymean"]), float(z["ystd"]))
            meta = json.loads(str(z["meta"]))
        self.feature_scales = int(meta.get("feature_scales", FEATURE_SCALES))
        return meta

    def close(self) -> None:
        if getattr(self, "_open", False):
            self._lib.VshipAMD_MLPFree(self.handle)
            self._open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- training data

SIZES = ((256, 256), (384, 216), (512, 288), (640, 360), (960, 540), (1280, 720), (1920, 1080))
SIZE_PROB = (0.15, 0.10, 0.10, 0.15, 0.15, 0.15, 0.20)
_D8 = synth._DCT8


def _dct_quantize_fast(plane: np.ndarray, q: float) -> np.ndarray:
    """8x8 orthonormal DCT-II quantisation with step q: the transform of
    synth.dct_quantize evaluated in float32 as whole-array matrix products
    (coefficients within float32 rounding of a quantisation tie may round the
    other way)."""
    h, w = plane.shape
    H, W = (h + 7) // 8 * 8, (w + 7) // 8 * 8
    p = np.pad(plane.astype(np.float32), ((0, H - h), (0, W - w)), mode="edge")
    d = _D8.astype(np.float32)
    a = (p.reshape(-1, 8) @ d.T).reshape(H // 8, 8, W // 8, 8)        # rows: [by, y, bx, kx]
    c = (a.transpose(0, 2, 3, 1).reshape(-1, 8) @ d.T)                 # cols: [by, bx, kx, ky]
    c = np.round(c / q) * q
    a = (c @ d).reshape(H // 8, W // 8, 8, 8).transpose(0, 3, 1, 2)   # back to [by, y, bx, kx]
    rec = (a.reshape(-1, 8) @ d).reshape(H, W)
    return rec[:h, :w]


def distort(planes, q: float, blur: float, noise: float, rng: np.random.Generator, noise_pool=None,
            chroma_resample: bool = False):
    """DCT quantisation (then, with chroma_resample, the chroma planes 2x box
    down- and replicated up-sampled as in synth.YuvSequence), then Gaussian
    blur, then Gaussian noise, per plane (chroma planes at their own
    resolution), rounded and clipped to 8 bits. With a noise pool
    (standard-normal float32 fields at least as large as the planes) the
    noise is a window of a pool field at a random offset."""
    out = []
    for c, p in enumerate(planes):
        v = _dct_quantize_fast(p, q)
        if chroma_resample and c > 0:
            v = synth._resample_chroma(v)
        if blur > 0.05:
            v = ndimage.gaussian_filter(v, blur, mode="reflect", truncate=3.0)
        if noise > 0:
            h, w = v.shape
            if noise_pool is not None:
                f = noise_pool[int(rng.integers(0, len(noise_pool)))]
                y0 = int(rng.integers(0, f.shape[0] - h + 1))
                x0 = int(rng.integers(0, f.shape[1] - w + 1))
                z = f[y0:y0 + h, x0:x0 + w]
            else:
                z = rng.standard_normal(v.shape, dtype=np.float32)
            v = v + z * np.float32(noise)
        out.append(np.clip(np.rint(v), 0, 255).astype(np.uint8))
    return out


class TrainingPairs:
    """Seeded source pool (sources_per_size independent frames per size) and
    per-pair distortion draws (see module doc)."""

    def __init__(self, seed: int = 0, sources_per_size: int = 16, sizes=SIZES, size_prob=SIZE_PROB):
        self.seed = seed
        self.sizes = sizes
        self.prob = np.asarray(size_prob, dtype=np.float64) / np.sum(size_prob)
        # every source is its own random sinusoid-sum / chroma-field draw
        self.pool = [[synth.YuvSequence(w, h, 1, seed=(seed * 1000 + k) * 1000 + j) for j in range(sources_per_size)]
                     for k, (w, h) in enumerate(sizes)]
        self.sources_per_size = sources_per_size
        nrng = np.random.default_rng([seed, 7])
        fh, fw = max(h for _, h in sizes) + 64, max(w for w, _ in sizes) + 64
        self.noise = [nrng.standard_normal((fh, fw), dtype=np.float32) for _ in range(3)]

    def pair(self, i: int):
        """(width, height, ref planes, dist planes, (q, blur, noise)) of pair i."""
        rng = np.random.default_rng([self.seed, i])
        k = int(rng.choice(len(self.sizes), p=self.prob))
        seq = self.pool[k][int(rng.integers(0, self.sources_per_size))]
        ref = self.source(seq, 0)
        q = float(np.exp(rng.uniform(np.log(2.0), np.log(64.0))))
        # a third of the pairs without blur and a third without noise, so pure
        # quantisation (the usual encoder candidate) is well covered
        blur = float(rng.uniform(0.0, 2.0)) if rng.uniform() >= 1 / 3 else 0.0
        noise = float(rng.uniform(0.0, 10.0)) if rng.uniform() >= 1 / 3 else 0.0
        resample = bool(rng.uniform() < 0.5)
        return seq.w, seq.h, ref, distort(ref, q, blur, noise, rng, self.noise, resample), (q, blur, noise)

    @staticmethod
    def source(seq, i):
        x0 = i * seq.shift
        cw, ch = (seq.w + 1) // 2, (seq.h + 1) // 2
        y = seq.luma[:, x0:x0 + seq.w]
        u = seq.chroma[0][:ch, x0 // 2:x0 // 2 + cw]
        v = seq.chroma[1][:ch, x0 // 2:x0 // 2 + cw]
        return [seq._quant(y), seq._quant(u), seq._quant(v)]


def label(pairs, batch: "S.SSIMU2Batch"):
    """Exact features (n, 108) and scores of host pairs (w, h, ref, dist, ...)
    through the batched metric (one call)."""
    align = lambda n: (n + 255) // 256 * 256
    total = sum(align(p.nbytes) for pr in pairs for p in pr[2] + pr[3])
    buf = S.DeviceBuffer(total)
    try:
        css, p1, p2, st = [], [], [], []
        off = 0
        for w, h, ref, dist, *_ in pairs:
            css.append(S.colorspace(w, h))
            trips = []
            for img in (ref, dist):
                trip = []
                for p in img:
                    buf.upload(p, off)
                    trip.append(buf.address + off)
                    off += align(p.nbytes)
                trips.append(trip)
            p1.append(trips[0])
            p2.append(trips[1])
            st.append(S.host_strides(ref))
        scores = batch.compute(css, css, p1, p2, st, st)
        feats = batch.features(len(pairs))
    finally:
        buf.free()
    return feats, scores


_GEN = None


def _pair_worker(i):
    return _GEN.pair(i)


def _worker_init():
    # one BLAS thread per worker process: the pool already fills the cores
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)


def generate(n: int, seed: int = 0, chunk: int = 1024, workers: int = 16, storage: int = S.STORAGE_FP32,
             log=None, processes: bool = False, pairs: "TrainingPairs | None" = None):
    """Features (n, 108), scores (n,), pixel counts (n,) and distortion
    parameters (n, 3) of n generated pairs, labelled by the exact metric.

    The pairs are synthesised by a pool (threads, or forked processes when
    `processes` is set: then this must run before the process makes any GPU
    call, since the workers are forked from it) while the GPU labels the
    previous chunk. `pairs` replaces the default TrainingPairs(seed)."""
    global _GEN
    _GEN = gen = pairs if pairs is not None else TrainingPairs(seed)
    feats = np.empty((n, 108), dtype=np.float64)
    scores = np.empty(n, dtype=np.float64)
    pixels = np.empty(n, dtype=np.float64)
    params = np.empty((n, 3), dtype=np.float64)
    t0 = time.perf_counter()
    if processes:
        import multiprocessing as mp
        pool = mp.get_context("fork").Pool(workers, initializer=_worker_init)
        submit = lambda idx: pool.map_async(_pair_worker, idx, chunksize=8)
        collect = lambda h: h.get()
    else:
        pool = cf.ThreadPoolExecutor(max_workers=workers)
        submit = lambda idx: [pool.submit(gen.pair, i) for i in idx]
        collect = lambda fs: [f.result() for f in fs]
    try:
        pending = submit(range(0, min(n, chunk)))
        bt = S.SSIMU2Batch(chunk, storage)
        for start in range(0, n, chunk):
            cur = collect(pending)
            pending = submit(range(start + chunk, min(n, start + 2 * chunk)))
            f, s = label(cur, bt)
            k = slice(start, start + len(cur))
            feats[k], scores[k] = f, s
            pixels[k] = [w * h for w, h, *_ in cur]
            params[k] = [pr[4] for pr in cur]
            if log:
                log(f"labelled {start + len(cur)}/{n} pairs ({time.perf_counter() - t0:.1f}s)")
        bt.close()
    finally:
        if processes:
            pool.terminate()
            pool.join()
        else:
            pool.shutdown()
    return feats, scores, pixels, params


# ---------------------------------------------------------------- screening

def screening_set(sources: int = 64, candidates: int = 16, width: int = 1920, height: int = 1080, seed: int = 5,
                  qmin: float = 2.0, qmax: float = 64.0, workers: int = 8):
    """Target-quality screening workload: `sources` YUV 4:2:0 8-bit frames and
    for each `candidates` versions with every plane 8x8 DCT-quantised at
    steps spaced geometrically from qmin to qmax
    (q_k = qmin * (qmax/qmin)^(k/(c-1))). Returns (refs, dists, qs) with
    dists[s][k] the planes of candidate k of source s."""
    qs = [qmin * (qmax / qmin) ** (k / max(candidates - 1, 1)) for k in range(candidates)]
    seq = synth.YuvSequence(width, height, sources, seed=seed, shift=16)
    refs = [TrainingPairs.source(seq, i) for i in range(sources)]

    def one(job):
        i, q = job
        return [np.clip(np.rint(_dct_quantize_fast(p, q)), 0, 255).astype(np.uint8) for p in refs[i]]

    from threadpoolctl import threadpool_limits
    with threadpool_limits(1), cf.ThreadPoolExecutor(max_workers=workers) as ex:
        flat = list(ex.map(one, [(i, q) for i in range(sources) for q in qs]))
    dists = [flat[i * candidates:(i + 1) * candidates] for i in range(sources)]
    return refs, dists, qs


def _select(scores, qs, target):
    """Highest-q candidate scoring at least `target`; the best-scoring one
    if none does."""
    ok = [k for k in range(len(qs)) if scores[k] >= target]
    return max(ok, key=lambda k: qs[k]) if ok else int(np.argmax(scores))


class Screening:
    """Device-resident screening workload and the two selection procedures."""

    def __init__(self, refs, dists, qs, width, height):
        self.qs = qs
        self.nsrc, self.ncand = len(refs), len(qs)
        self.cs = S.colorspace(width, height)
        self.strides = S.host_strides(refs[0])
        align = lambda n: (n + 255) // 256 * 256
        per = sum(align(p.nbytes) for p in refs[0])
        self.buf = S.DeviceBuffer(per * self.nsrc * (1 + self.ncand))
        off = 0
        self.ref_ptr, self.cand_ptr = [], []
        for i in range(self.nsrc):
            trip = []
            for p in refs[i]:
                self.buf.upload(p, off)
                trip.append(self.buf.address + off)
                off += align(p.nbytes)
            self.ref_ptr.append(trip)
            row = []
            for k in range(self.ncand):
                trip = []
                for p in dists[i][k]:
                    self.buf.upload(p, off)
                    trip.append(self.buf.address + off)
                    off += align(p.nbytes)
                row.append(trip)
            self.cand_ptr.append(row)
        self.pixels = width * height
        self._prepared = {}  # call arguments of the full candidate list

    def _run(self, batch, items, scale_mask=S.ALL_SCALES, features=True):
        """Scores and features of (source, candidate) items in one batched
        call (scale_mask < ALL_SCALES: only those scales blurred, see
        SSIMU2Batch.complete_scales). The source planes are passed once per
        source (the same device pointers), so the batch computes each source's
        pyramid once; the call arguments of the full candidate list are built
        once."""
        key = tuple(items)
        args = self._prepared.get(key)
        if args is None:
            css = [self.cs] * len(items)
            p1 = [self.ref_ptr[i] for i, _ in items]
            p2 = [self.cand_ptr[i][k] for i, k in items]
            st = [self.strides] * len(items)
            args = batch.prepare(css, css, p1, p2, st, st)
            if len(items) == self.nsrc * self.ncand:
                self._prepared[key] = args
        scores = batch.compute_prepared(args, np.zeros(len(items), dtype=np.float64), scale_mask)
        return scores, (batch.features(len(items)) if features else None)

    def exact_only(self, batch, target):
        items = [(i, k) for i in range(self.nsrc) for k in range(self.ncand)]
        scores, _ = self._run(batch, items, features=False)
        sc = scores.reshape(self.nsrc, self.ncand)
        return [_select(sc[i], self.qs, target) for i in range(self.nsrc)], sc

    def with_surrogate(self, batch, net, target, top=4, fallback_sigma=1.5):
        """Surrogate-screened selection. Returns (choices, stats, mean, sigma).

        One batched call blurs the scales of the network's inputs for every
        candidate (the planes of the other scales are stored on the way). The network
        ranks every candidate; per source the `top` highest-q candidates
        predicted to reach the target (margin: two predicted standard
        deviations) plus every candidate whose predicted standard deviation
        exceeds fallback_sigma are confirmed by the exact metric, which only
        has the other scales left to blur for them (their exact scores equal a full
        evaluation's bit for bit); a source whose shortlist has no candidate
        reaching the target is completed for all its candidates and selected
        exactly. stats["confirmed"] maps (source, candidate) to the exact
        scores used."""
        items = [(i, k) for i in range(self.nsrc) for k in range(self.ncand)]
        scales = net.feature_scales
        rest_scales = S.ALL_SCALES & ~scales
        t0 = time.perf_counter()
        _, feats = self._run(batch, items, scales)
        t1 = time.perf_counter()
        X = inputs_from_features(feats, np.full(len(items), float(self.pixels)), scales)
        mean, sigma = net.predict(X)
        t2 = time.perf_counter()
        mean = mean.reshape(self.nsrc, self.ncand)
        sigma = sigma.reshape(self.nsrc, self.ncand)
        short = []
        fallback_items = 0
        for i in range(self.nsrc):
            order = sorted(range(self.ncand), key=lambda k: -self.qs[k])
            likely = [k for k in order if mean[i, k] + 2.0 * sigma[i, k] >= target][:top]
            if not likely:
                likely = sorted(range(self.ncand), key=lambda k: -mean[i, k])[:top]
            unsure = [k for k in range(self.ncand) if sigma[i, k] > fallback_sigma and k not in likely]
            fallback_items += len(unsure)
            short.append(sorted(set(likely) | set(unsure)))
        confirm = [(i, k) for i in range(self.nsrc) for k in short[i]]
        cscores = batch.complete_scales([i * self.ncand + k for i, k in confirm], rest_scales)
        exact = dict(zip(confirm, cscores))
        choices, full_fallback = [], []
        for i in range(self.nsrc):
            ok = [k for k in short[i] if exact[(i, k)] >= target]
            if ok:
                choices.append(max(ok, key=lambda k: self.qs[k]))
            else:
                full_fallback.append(i)
                choices.append(None)
        if full_fallback:
            rest = [(i, k) for i in full_fallback for k in range(self.ncand) if (i, k) not in exact]
            if rest:
                rs = batch.complete_scales([i * self.ncand + k for i, k in rest], rest_scales)
                exact.update(zip(rest, rs))
            for i in full_fallback:
                choices[i] = _select([exact[(i, k)] for k in range(self.ncand)], self.qs, target)
        t3 = time.perf_counter()
        stats = {
            "feature_pass_s": t1 - t0, "network_s": t2 - t1, "confirm_s": t3 - t2, "total_s": t3 - t0,
            "confirmed_pairs": len(confirm), "uncertain_pairs_sent_to_exact": fallback_items,
            "sources_fully_rescored": len(full_fallback), "confirmed": exact,
        }
        return choices, stats, mean, sigma

    def close(self):
        self.buf.free()


def evaluate(net: SurrogateNet, X, y, fallback_sigma: float):
    mean, sigma = net.predict(X)
    err = np.abs(mean - y)
    fb = sigma > fallback_sigma
    after = np.where(fb, 0.0, err)
    return {
        "n": int(len(y)),
        "mae": float(err.mean()),
        "p99_abs_error": float(np.percentile(err, 99)),
        "max_abs_error": float(err.max()),
        "fallback_sigma": fallback_sigma,
        "fallback_rate": float(fb.mean()),
        "mae_after_fallback": float(after.mean()),
        "p99_after_fallback": float(np.percentile(after, 99)),
        "sigma_error_corr": float(np.corrcoef(sigma, err)[0, 1]),
    }


def linear_baseline(Xtr, ytr, Xte, yte):
    """Least-squares affine map from the same 55 inputs (the method without
    the network's non-linearity)."""
    A = np.hstack([Xtr.astype(np.float64), np.ones((len(Xtr), 1))])
    coef, *_ = np.linalg.lstsq(A, ytr, rcond=None)
    pred = np.hstack([Xte.astype(np.float64), np.ones((len(Xte), 1))]) @ coef
    err = np.abs(pred - yte)
    return {"mae": float(err.mean()), "p99_abs_error": float(np.percentile(err, 99))}


def _train_main(argv):
    ap = argparse.ArgumentParser(prog="python -m vship_amd.surrogate train")
    ap.add_argument("--pairs", type=int, default=200000)
    ap.add_argument("--holdout", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--fallback-sigma", type=float, default=1.5)
    ap.add_argument("--feature-scales", type=lambda v: int(v, 0), default=FEATURE_SCALES,
                    help="scale mask of the inputs: 0b00111 (scales 0-2) or 0b11100 (scales 2-4)")
    ap.add_argument("--out", default=None, help="default: data/surrogate.npz or data/surrogate_coarse.npz")
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    feature_index(a.feature_scales)  # three scales
    if a.out is None:
        a.out = PARAMS_FILE if a.feature_scales == FEATURE_SCALES else COARSE_PARAMS_FILE
    log = lambda m: print(m, flush=True)
    t0 = time.perf_counter()
    feats, scores, pixels, dparams = generate(a.pairs, a.seed, workers=a.workers, log=log, processes=True)
    t_data = time.perf_counter() - t0
    X = inputs_from_features(feats, pixels, a.feature_scales)
    y = scores.astype(np.float32)
    ntr = a.pairs - a.holdout
    with SurrogateNet(max_batch=a.batch) as net:
        net.feature_scales = a.feature_scales
        net.fit_norm(X[:ntr], y[:ntr])
        net.set_params(init_params(a.seed))
        t1 = time.perf_counter()
        losses = net.train(X[:ntr], y[:ntr], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
        t_train = time.perf_counter() - t1
        log(f"trained {a.epochs} epochs on {ntr} pairs in {t_train:.2f}s; loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        test = evaluate(net, X[ntr:], y[ntr:], a.fallback_sigma)
        train = evaluate(net, X[:ntr][:20000], y[:ntr][:20000], a.fallback_sigma)
        report = {
            "pairs": a.pairs, "train_pairs": ntr, "holdout_pairs": a.holdout, "epochs": a.epochs,
            "batch": a.batch, "lr": a.lr, "seed": a.seed, "feature_scales": a.feature_scales,
            "data_seconds": round(t_data, 2), "train_seconds": round(t_train, 3),
            "loss_per_epoch": [round(float(v), 5) for v in losses],
            "holdout": test, "train_subset": train,
            "linear_baseline_holdout": linear_baseline(X[:ntr], y[:ntr], X[ntr:], y[ntr:]),
            "score_range": [float(y.min()), float(y.max())], "score_mean": float(y.mean()),
        }
        net.save(a.out, pairs=a.pairs, epochs=a.epochs, seed=a.seed, holdout_mae=test["mae"],
                 fallback_sigma=a.fallback_sigma)
    log(json.dumps(report))
    if a.report:
        os.makedirs(os.path.dirname(os.path.abspath(a.report)), exist_ok=True)
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return 0


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] != "train":
        print("usage: python -m vship_amd.surrogate train [--pairs N] [--epochs E] [--feature-scales MASK] "
              "[--out FILE] [--report FILE]")
        return 2
    return _train_main(argv[1:])


if __name__ == "__main__":
    sys.exit(main())
