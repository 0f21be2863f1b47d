"""Fixed workloads for profiling the SSIMULACRA2 kernels (rocprofv3 targets).

python -m vship_amd.profile_workload plain --pairs 8     # plain FP32 path, 1080p YUV420
python -m vship_amd.profile_workload batch --pairs 240   # batched FP16 path, 1080p YUV420
python -m vship_amd.profile_workload batch --blur-config 5  # same, matrix-core blur
python -m vship_amd.profile_workload batch --sources 16  # 240 pairs over 16 shared reference images
python -m vship_amd.profile_workload mixed --pairs 512   # 4K 10-bit / 1080p / 720p / 360p in one batch
python -m vship_amd.profile_workload copy --mib 256      # plain copy kernel (counter calibration)
python -m vship_amd.profile_workload pcie --handlers 8 --pairs 16 --reps 15
    # Vship_ComputeSSIMU2 on pinned host planes, one host thread per handler

Each mode prints what it ran and the HIP-event time per kernel.
"""
from __future__ import annotations

import argparse
import ctypes
import sys
import threading
import time

import numpy as np

from . import ssimu2 as S
from . import synth


def _upload_pairs(n: int, seed: int = 0):
    seq = synth.YuvSequence(1920, 1080, n, seed=seed)
    bufs, ptrs = [], []
    strides = None
    for i in range(n):
        ref, dist, _ = seq.frame(i)
        strides = S.host_strides(ref)
        pair = []
        for img in (ref, dist):
            trip = []
            for p in img:
                b = S.DeviceBuffer(p.nbytes)
                b.upload(p)
                bufs.append(b)
                trip.append(b.address)
            pair.append(trip)
        ptrs.append(pair)
    return bufs, ptrs, strides


def upload_mixed(pairs):
    """Planes of mixed-size pairs (synth.mixed_batch) in one device buffer
    (256-byte aligned planes); returns (buf, colorspaces, ptrs[pair][img][3],
    strides[pair][3])."""
    align = lambda n: (n + 255) // 256 * 256
    total = sum(align(p.nbytes) for *_, ref, dist, _q in pairs for p in ref + dist)
    buf = S.DeviceBuffer(total)
    css, ptrs, strides = [], [], []
    off = 0
    for w, h, bits, ref, dist, _q in pairs:
        css.append(S.colorspace(w, h, sample=S.SAMPLE_UINT10 if bits == 10 else S.SAMPLE_UINT8))
        pair = []
        for img in (ref, dist):
            planes = []
            for p in img:
                buf.upload(p, off)
                planes.append(buf.address + off)
                off += align(p.nbytes)
            pair.append(planes)
        ptrs.append(pair)
        strides.append(S.host_strides(ref))
    return buf, css, ptrs, strides


def _mixed(a) -> int:
    """a.pairs pairs of the mixed sizes (3840x2160 10-bit, 1080p, 720p, 360p
    8-bit YUV420, split 1:4:8:16, seed 3) in one batched call, a.reps calls."""
    counts = synth.mixed_counts(a.pairs)
    pairs = synth.mixed_batch(counts, seed=3, workers=16, processes=True)  # forks: before the first HIP call
    buf, css, ptrs, strides = upload_mixed(pairs)
    print(f"mixed: {a.pairs} pairs {counts} (3840x2160 10-bit / 1920x1080 / 1280x720 / 640x360 8-bit), "
          f"{a.reps} reps", flush=True)
    storage = S.STORAGE_FP16 if a.storage == "fp16" else S.STORAGE_FP32
    try:
        with S.SSIMU2Batch(a.pairs, storage) as b:
            args = (css, css, [p[0] for p in ptrs], [p[1] for p in ptrs], strides, strides)
            for _ in range(a.reps):
                scores = b.compute(*args)
            fr, bl, sc, gr = b.profile(a.reps)
            print(f"tuning {b.tuning()}; scores mean {np.mean(scores):.5f}; per batch ms frontend {fr:.4f} "
                  f"blur {bl:.4f} score {sc:.4f} graph {gr:.4f}", flush=True)
    finally:
        buf.free()
    return 0


def _pcie(a) -> int:
    """a.pairs frame pairs in pinned host memory, a.reps passes over them by
    a.handlers handlers (pair i on handler i mod n), one host thread each."""
    lib = S.load_library()
    seq = synth.YuvSequence(1920, 1080, a.pairs, seed=0)
    pinned, pairs, strides = [], [], None
    nbytes = 0
    for i in range(a.pairs):
        ref, dist, _ = seq.frame(i)
        strides = S.host_strides(ref)
        pair = []
        for img in (ref, dist):
            trip = []
            for p in img:
                ptr = ctypes.c_void_p()
                S.check(lib.Vship_PinnedMalloc(ctypes.byref(ptr), p.nbytes))
                ctypes.memmove(ptr, p.ctypes.data, p.nbytes)
                pinned.append(ptr)
                trip.append(ptr.value)
                nbytes += p.nbytes if i == 0 else 0
            pair.append(trip)
        pairs.append((S._plane_args(pair[0], strides), S._plane_args(pair[1], strides)))
    cs = S.colorspace(1920, 1080)
    n = a.handlers
    handlers = [S.SSIMU2Handler(cs, cs) for _ in range(n)]
    total = a.pairs * a.reps

    def worker(k, count):
        sc = ctypes.c_double()
        for i in range(k, count, n):
            (p1, l1), (p2, l2) = pairs[i % len(pairs)]
            S.check(lib.Vship_ComputeSSIMU2(handlers[k].handle, ctypes.byref(sc), p1, p2, l1, l2))

    def run(count):
        ths = [threading.Thread(target=worker, args=(k, count)) for k in range(n)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()

    try:
        run(2 * n)  # each handler's first host-plane call allocates its staging buffers
        t0 = time.perf_counter()
        run(total)
        dt = time.perf_counter() - t0
    finally:
        for h in handlers:
            h.close()
        for p in pinned:
            lib.Vship_PinnedFree(p)
    print(f"pcie: {n} handlers, {total} pairs of pinned host planes: {total / dt:.1f} pairs/s, "
          f"{total * nbytes / dt / 1e9:.2f} GB/s host to device", flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["plain", "batch", "mixed", "copy", "pcie"])
    ap.add_argument("--handlers", type=int, default=8, help="pcie: handlers (one host thread each)")
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--storage", choices=["fp16", "fp32"], default="fp16")
    ap.add_argument("--blur-config", type=int, default=None, help="batch: blur tile config (5-8 = matrix cores)")
    ap.add_argument("--blur-waves", type=int, default=None, help="batch: blur workgroups per CU")
    ap.add_argument("--front-waves", type=int, default=None, help="batch: front-end workgroups per CU")
    ap.add_argument("--sources", type=int, default=0,
                    help="batch: pairs share the reference planes of this many sources, in contiguous groups "
                         "(0: every pair has its own reference)")
    a = ap.parse_args(argv)
    if a.mode == "copy":
        nbytes = a.mib << 20
        src, dst = S.DeviceBuffer(nbytes), S.DeviceBuffer(nbytes)
        ms = ctypes.c_double()
        S.check(S.load_library().VshipAMD_DeviceCopy(ctypes.c_void_p(dst.address), ctypes.c_void_p(src.address),
                                                       nbytes, a.reps, ctypes.byref(ms)))
        print(f"copy {nbytes} bytes: {ms.value:.4f} ms, {2 * nbytes / ms.value / 1e6:.1f} GB/s read+write", flush=True)
        return 0
    if a.mode == "pcie":
        return _pcie(a)
    if a.mode == "mixed":
        return _mixed(a)
    bufs, ptrs, st = _upload_pairs(a.pairs)
    cs = S.colorspace(1920, 1080)
    print(f"{a.mode}: {a.pairs} pairs 1920x1080 yuv420p8, {a.reps} reps", flush=True)
    if a.mode == "plain":
        with S.SSIMU2Handler(cs, cs) as h:
            for _ in range(a.reps):
                scores = [h.compute_device(p[0], p[1], st, st) for p in ptrs]
            fr, bl, sc = h.profile(a.reps)
        print(f"scores mean {np.mean(scores):.5f}; per pair ms frontend {fr:.4f} blur {bl:.4f} score {sc:.4f}",
              flush=True)
    else:
        storage = S.STORAGE_FP16 if a.storage == "fp16" else S.STORAGE_FP32
        with S.SSIMU2Batch(a.pairs, storage) as b:
            if a.blur_config is not None or a.blur_waves is not None or a.front_waves is not None:
                t = b.tuning()
                b.set_tuning(t["blur_config"] if a.blur_config is None else a.blur_config,
                             t["blur_waves"] if a.blur_waves is None else a.blur_waves,
                             t["front_waves"] if a.front_waves is None else a.front_waves)
            group = -(-a.pairs // a.sources) if a.sources > 0 else 1
            refs = [ptrs[i // group * group][0] for i in range(a.pairs)]  # same device planes within a group
            args = ([cs] * a.pairs, [cs] * a.pairs, refs, [p[1] for p in ptrs], [st] * a.pairs, [st] * a.pairs)
            for _ in range(a.reps):
                scores = b.compute(*args)
            fr, bl, sc, gr = b.profile(a.reps)
            print(f"tuning {b.tuning()}; scores mean {np.mean(scores):.5f}; per batch ms frontend {fr:.4f} "
                  f"blur {bl:.4f} score {sc:.4f} graph {gr:.4f}", flush=True)
    for b in bufs:
        b.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
