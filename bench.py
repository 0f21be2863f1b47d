"""SSIMULACRA2 throughput on one MI355X: 240 synthetic 1080p YUV420 frame pairs.

A step is one pass of the plain FP32 path over the 240 pairs (8-bit YUV 4:2:0
BT.709 limited, translating sinusoid content, distorted by per-frame 8x8 DCT
quantisation with q ~ U{4..32} and chroma re-subsampling). Inputs are resident
in HBM before the timed region; the pairs are spread over --handlers handlers,
one host thread each, every pair a blocking VshipAMD_ComputeSSIMU2Device call
(the drop-in path with device-resident planes).

Prints one JSON line (rank 0) with the throughput (also at HIP's default of
4 hardware queues, from a child process), the roofline of the dominant kernel
(HIP-event timed on its own stream after the timed region: VALU issue rate
from its counted instructions, and HBM rate from its algorithmic bytes), the
batched FP16, mixed-size and screening legs, the PCIe-inclusive rate of
Vship_ComputeSSIMU2 on pinned host planes with 4 and 8 handlers, and the FP64
oracle's CPU baseline on bounded samples (one core and the box's share).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes
import json
import os
import subprocess
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from vship_amd import ssimu2 as S  # noqa: E402
from vship_amd import synth  # noqa: E402
from vship_amd.profile_workload import upload_mixed  # noqa: E402

W, H, PAIRS = 1920, 1080, 240
HBM_PEAK_GBS = 8000.0
FRONT_KERNEL = "frontend_single_kernel<float, 1>"  # the plain path's front-end, 8-bit YUV 4:2:0 BT.709 instantiation
BLUR_KERNEL = "blur_strip_kernel<float>"  # the plain path's blur/map kernel (matrix cores, 64x128 strips)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores() -> int:
    """Cores this process may run on (the box's CPU share may be smaller)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# worker processes of the synthesis (forked before the first HIP call), and
# the cap on oracle threads: one GPU's share of the box's CPUs
WORKERS = max(1, min(16, host_cores()))


def make_frames(n: int, seed: int = 0):
    """(ref, dist, q) of the n pairs, generated in forked worker processes."""
    seq = synth.YuvSequence(W, H, n, bits=8, seed=seed)
    return [f[3:] for f in synth.frames_in_processes({0: seq}, [(0, i) for i in range(n)], WORKERS)]


def upload(frames):
    """One device buffer holding every plane; returns (buf, ptrs[pair][img][3], strides)."""
    sizes = [p.nbytes for p in frames[0][0]]
    per_img = sum(sizes)
    buf = S.DeviceBuffer(per_img * 2 * len(frames))
    ptrs = []
    off = 0
    for ref, dist, _ in frames:
        pair = []
        for img in (ref, dist):
            planes = []
            for p in img:
                buf.upload(p, off)
                planes.append(buf.address + off)
                off += p.nbytes
            pair.append(planes)
        ptrs.append(pair)
    strides = [int(p.strides[0]) for p in frames[0][0]]
    return buf, ptrs, strides


def plain_args(ptrs, strides):
    """ctypes arguments of every pair's VshipAMD_ComputeSSIMU2Device call, built once (as a C caller
    holds its plane arrays), so that the timed loop is the library call itself."""
    out = []
    for pair in ptrs:
        p1, l1 = S._plane_args(pair[0], strides)
        p2, l2 = S._plane_args(pair[1], strides)
        out.append((p1, l1, p2, l2))
    return out


def run_plain(handlers, args, scores):
    n = len(handlers)
    errors = []
    lib = S.load_library()

    def worker(k):
        try:
            h = handlers[k].handle
            sc = ctypes.c_double()
            ref = ctypes.byref(sc)
            call = lib.VshipAMD_ComputeSSIMU2Device
            for i in range(k, len(args), n):
                p1, l1, p2, l2 = args[i]
                rc = call(h, ref, p1, p2, l1, l2)
                if rc != 0:
                    S.check(rc)
                scores[i] = sc.value
        except Exception as e:  # surfaced after join
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(n)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]


def make_mixed(npairs: int):
    """The mixed-size pairs (forked workers: call before the first HIP call)."""
    counts = synth.mixed_counts(npairs)
    t0 = time.perf_counter()
    pairs = synth.mixed_batch(counts, seed=3, workers=WORKERS, processes=True)
    log(f"mixed batch: generated {npairs} pairs {counts} in {time.perf_counter() - t0:.1f}s")
    return counts, pairs


def oracle_sample(pairs, per_size: int = 1):
    """Indices of a stratified sample: the first per_size pairs of each size."""
    seen, idx = {}, []
    for i, (w, h, bits, *_rest) in enumerate(pairs):
        if seen.get((w, h, bits), 0) < per_size:
            seen[(w, h, bits)] = seen.get((w, h, bits), 0) + 1
            idx.append(i)
    return idx


def mixed_leg(counts, pairs, steps: int, warmup: int):
    """Uneven sizes in one batched FP16 call: 3840x2160 10-bit, 1080p, 720p and
    360p 8-bit YUV420 pairs split 1:4:8:16, seeded shuffled order. The FP16
    scores of a stratified sample (one pair of each size) are compared with
    the FP64 oracle."""
    import oracle
    npairs = len(pairs)
    buf, css, ptrs, strides = upload_mixed(pairs)
    p1 = [p[0] for p in ptrs]
    p2 = [p[1] for p in ptrs]
    px = sum(w * h for w, h, *_ in pairs)
    try:
        with S.SSIMU2Batch(npairs, S.STORAGE_FP16) as bt:
            tuning, tune_ms = bt.autotune(css, css, p1, p2, strides, strides, reps=2)
            prepared = bt.prepare(css, css, p1, p2, strides, strides)
            fp16 = np.zeros(npairs, dtype=np.float64)
            for _ in range(warmup):
                bt.compute_prepared(prepared, fp16)
            S.check(S.load_library().VshipAMD_DeviceSynchronize())
            t0 = time.perf_counter()
            for _ in range(steps):
                bt.compute_prepared(prepared, fp16)
            dt = (time.perf_counter() - t0) / steps
            fr, bl, sc, gr = bt.profile(3)
            bbytes, bflops = bt.work()
        with S.SSIMU2Batch(npairs, S.STORAGE_FP32) as bt:
            fp32 = bt.compute(css, css, p1, p2, strides, strides)
    finally:
        buf.free()
    value = npairs / dt
    eq1080 = px / (1920 * 1080)
    log(f"mixed batch fp16: {value:.1f} pairs/s = {eq1080 / dt:.1f} 1080p-equivalent pairs/s, tuning {tuning}")
    sample = oracle_sample(pairs)

    def exact(i):
        *_wh, ref, dist, _q = pairs[i]
        return oracle.ssimu2(css[i], ref, strides[i], css[i], dist, strides[i])[0]

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=min(WORKERS, len(sample))) as ex:  # C calls release the GIL
        want = list(ex.map(exact, sample))
    vs_oracle = [{"size": f"{pairs[i][0]}x{pairs[i][1]} {pairs[i][2]}-bit", "fp16": round(float(fp16[i]), 5),
                  "fp32": round(float(fp32[i]), 5), "oracle": round(w, 5),
                  "abs_delta_fp16": round(abs(float(fp16[i]) - w), 6)} for i, w in zip(sample, want)]
    log(f"mixed batch: FP64 oracle on {len(sample)} pairs in {time.perf_counter() - t0:.1f}s, "
        f"max |fp16 - oracle| {max(v['abs_delta_fp16'] for v in vs_oracle):.5f}")
    return {
        "workload": f"{npairs} pairs YUV420 BT.709 limited, 3840x2160 10-bit / 1920x1080 / 1280x720 / 640x360 8-bit "
                    f"= {counts}, one batched FP16-storage call (one graph replay)",
        "value": round(value, 2), "unit": "pairs/s", "ms_per_batch": round(dt * 1e3, 3),
        "equiv_1080p_pairs_per_s": round(eq1080 / dt, 2),
        "tuning": tuning,
        "kernel_ms_per_batch": {"frontend": round(fr, 4), "blur": round(bl, 4), "score": round(sc, 4),
                                "graph": round(gr, 4)},
        "path_gbs": round(bbytes / (gr * 1e-3) / 1e9, 1),
        "max_abs_delta_fp16_vs_fp32": round(float(np.max(np.abs(fp16 - fp32))), 6),
        "max_abs_delta_fp16_vs_oracle": max(v["abs_delta_fp16"] for v in vs_oracle),
        "oracle_sample": vs_oracle,
    }


def screening_leg(steps: int, target: float = 70.0):
    """Target-quality screening: 64 1080p YUV420 sources x 16 DCT candidates
    (q from 2 to 64, geometric). Exact-only selection (every candidate
    through the batched FP16 metric) against surrogate screening: one pass
    over the scales of the network's inputs for every candidate, exact
    confirmation of the top 4 per source by completing their other scales,
    exact fallback for uncertain candidates. The headline network reads the
    scale 2-4 features (6% of the blur); the network on the scale 0-2
    features is reported beside it."""
    from vship_amd import surrogate as SG
    nets = [("coarse", SG.COARSE_PARAMS_FILE), ("scales012", SG.PARAMS_FILE)]
    missing = [f for _, f in nets if not os.path.exists(f)]
    if missing:
        return {"skipped": f"no trained surrogate at {', '.join(os.path.relpath(f, ROOT) for f in missing)}"}
    t0 = time.perf_counter()
    refs, dists, qs = SG.screening_set(64, 16, workers=16)
    log(f"screening: generated 64 x 16 candidates in {time.perf_counter() - t0:.1f}s")
    scr = SG.Screening(refs, dists, qs, 1920, 1080)
    ncand = scr.nsrc * scr.ncand
    runs = {}
    try:
        with S.SSIMU2Batch(ncand, S.STORAGE_FP16) as bt:
            scr.exact_only(bt, target)
            t0 = time.perf_counter()
            for _ in range(steps):
                exact_choice, exact_scores = scr.exact_only(bt, target)
            t_exact = (time.perf_counter() - t0) / steps
            for name, path in nets:
                with SG.SurrogateNet(max_batch=1024) as net:
                    meta = net.load(path)
                    scr.with_surrogate(bt, net, target)
                    t0 = time.perf_counter()
                    for _ in range(steps):
                        choice, stats, mean, sigma = scr.with_surrogate(bt, net, target)
                    runs[name] = ((time.perf_counter() - t0) / steps, choice, stats, mean, sigma,
                                  dict(meta, feature_scales=net.feature_scales))
    finally:
        scr.close()

    def report(name):
        t_sur, choice, stats, mean, sigma, meta = runs[name]
        err = np.abs(mean - exact_scores)
        return {
            "feature_scales": sorted(s for s in range(5) if meta["feature_scales"] >> s & 1),
            "surrogate_candidates_per_s": round(ncand / t_sur, 1),
            "speedup_vs_exact_only": round(t_exact / t_sur, 3),
            "agreement_with_exact_only": float(np.mean([a == b for a, b in zip(exact_choice, choice)])),
            "confirmed_scores_equal_exact_only": all(v == exact_scores[i, k]
                                                     for (i, k), v in stats["confirmed"].items()),
            "confirmed_pairs": stats["confirmed_pairs"],
            "uncertain_pairs_sent_to_exact": stats["uncertain_pairs_sent_to_exact"],
            "sources_fully_rescored": stats["sources_fully_rescored"],
            "seconds": {"feature_pass": round(stats["feature_pass_s"], 4), "network": round(stats["network_s"], 5),
                        "confirm": round(stats["confirm_s"], 4)},
            "prediction_mae": round(float(err.mean()), 4),
            "prediction_p99": round(float(np.percentile(err, 99)), 4),
            "fallback_rate": round(float(np.mean(sigma > meta.get("fallback_sigma", 1.5))), 4),
            "model": meta,
        }

    head, spec = report("coarse"), report("scales012")
    log(f"screening: exact-only {ncand / t_exact:.0f} candidates/s, surrogate (scales 2-4) "
        f"{head['surrogate_candidates_per_s']:.0f} candidates/s agreement {head['agreement_with_exact_only']:.3f}, "
        f"surrogate (scales 0-2) {spec['surrogate_candidates_per_s']:.0f} agreement "
        f"{spec['agreement_with_exact_only']:.3f}")
    out = {
        "workload": "64 sources 1920x1080 YUV420 8-bit x 16 candidates (8x8 DCT quantisation, q = 2 * 32^(k/15)), "
                    f"target score {target}: pick the highest-q candidate reaching it",
        "exact_only_candidates_per_s": round(ncand / t_exact, 1),
        "exact_only_seconds": round(t_exact, 4),
    }
    out.update(head)
    out["scales012_network"] = spec
    return out


def cpu_baseline(frames, threads: int, pairs: int):
    import oracle
    cs = S.colorspace(W, H)
    sample = [frames[i % len(frames)] for i in range(pairs)]

    def one(fr):
        ref, dist, _ = fr
        st = S.host_strides(ref)
        return oracle.ssimu2(cs, ref, st, cs, dist, st)[0]

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:  # the C oracle releases the GIL
        vals = list(ex.map(one, sample))
    dt = time.perf_counter() - t0
    return pairs / dt, vals


# the plain path's kernels and their counters (rocprofv3 runs summarised by
# vship_amd.profile_summary counters into this file)
COUNTERS_FILE = os.path.join(ROOT, "profiles", "kernel_counters_1080p.json")
# VALU issue peak: 1024 SIMDs x 16 lanes x 2.4 GHz (a wave64 VALU instruction
# holds a SIMD for 4 cycles; packed FP32 is disabled in this build)
VALU_LANE_OPS_PEAK = 1024 * 16 * 2.4e9


def load_counters():
    if not os.path.exists(COUNTERS_FILE):
        return {}
    return json.load(open(COUNTERS_FILE))


def valu_roof(counters, workload: str, kernel: str, ms: float, pairs: float):
    """VALU issue rate of one launch: the kernel's VALU instructions per pair
    (counters of the same kernel and workload) over its HIP-event time."""
    k = counters.get(workload, {}).get("kernels", {}).get(kernel)
    if not k or "valu_insts_per_pair" not in k:
        return None
    lane_ops = k["valu_insts_per_pair"] * 64.0 * pairs
    achieved = lane_ops / (ms * 1e-3)
    return {"lane_ops_per_launch": lane_ops, "achieved_tops": round(achieved / 1e12, 3),
            "peak_tops": round(VALU_LANE_OPS_PEAK / 1e12, 3), "frac": round(achieved / VALU_LANE_OPS_PEAK, 4),
            "valu_busy_profiled": round(k.get("valu_busy", float("nan")), 4),
            "mfma_busy_profiled": round(k.get("mfma_busy", float("nan")), 4)}


def path_valu(counters, workload: str, kernels, pairs_per_s: float):
    """VALU issue rate of the whole path over its timed region: the counted
    VALU instructions of a pair's kernels times the pairs/s, against the peak
    (a single pair's launch does not fill the GPU; the handlers' kernels
    together do)."""
    ks = counters.get(workload, {}).get("kernels", {})
    if not all(k in ks and "valu_insts_per_pair" in ks[k] for k in kernels):
        return None
    lane_ops = sum(ks[k]["valu_insts_per_pair"] for k in kernels) * 64.0
    achieved = lane_ops * pairs_per_s
    return {"kernels": list(kernels), "lane_ops_per_pair": lane_ops, "achieved_tops": round(achieved / 1e12, 3),
            "peak_tops": round(VALU_LANE_OPS_PEAK / 1e12, 3), "frac": round(achieved / VALU_LANE_OPS_PEAK, 4)}


def run_plain_timed(handlers, call_args, scores, steps: int, warmup: int):
    for _ in range(warmup):
        run_plain(handlers, call_args, scores)
    lib = S.load_library()
    S.check(lib.VshipAMD_DeviceSynchronize())
    t0 = time.perf_counter()
    for _ in range(steps):
        run_plain(handlers, call_args, scores)
    S.check(lib.VshipAMD_DeviceSynchronize())
    return time.perf_counter() - t0


DRIVER = os.path.join(ROOT, "vship_amd", "ffvship_amd")


def driver_leg(frames, gpu_threads=(4, 8), reader_threads: int = 4):
    """The throughput driver (FFVship's pipeline over raw YUV files: reader
    threads fill pinned buffers, -g handler threads call Vship_ComputeSSIMU2)
    on the 240 pairs written as two page-cached raw 1080p files. Runs before
    this process initialises HIP; the driver is a child process."""
    if not os.path.exists(DRIVER):
        return {"skipped": f"{os.path.relpath(DRIVER, ROOT)} is not built"}
    import tempfile
    out = {"workload": f"{len(frames)} frame pairs 1920x1080 yuv420p (raw files, page cache), -t {reader_threads}",
           "reference_context": "FFVship, README.md:204: 1339 frames in 9 s (148.8 fps) including video decode"}
    with tempfile.TemporaryDirectory(prefix="vship_bench_") as d:
        src, enc = os.path.join(d, "s.yuv"), os.path.join(d, "e.yuv")
        with open(src, "wb") as fs, open(enc, "wb") as fe:
            for ref, dist, _ in frames:
                for p in ref:
                    fs.write(np.ascontiguousarray(p).tobytes())
                for p in dist:
                    fe.write(np.ascontiguousarray(p).tobytes())
        for g in gpu_threads:
            cmd = [DRIVER, "--source", src, "--encoded", enc, "--width", str(W), "--height", str(H),
                   "-g", str(g), "-t", str(reader_threads)]
            r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=300)
            if r.returncode != 0:
                raise RuntimeError(f"driver run failed with {r.returncode}")
            fps = [float(l.split(" at ")[1].split()[0]) for l in r.stdout.splitlines() if l.startswith("Computed")]
            out[f"fps_g{g}"] = fps[0] if fps else None
            log(f"driver -g {g} -t {reader_threads}: {out[f'fps_g{g}']} fps")
    return out


def pcie_leg(frames, strides, cs, handler_counts=(4, 8), n_host: int = 240):
    """Vship_ComputeSSIMU2 (the reference's call: planes in pinned host memory,
    copied to the device inside the call) with 4 and 8 handlers, one host
    thread each: pairs/s and the achieved host-to-device rate."""
    lib = S.load_library()
    pinned = []
    out = {}
    try:
        host_pairs = []
        for ref, dist, _ in frames[:16]:
            pair = []
            for img in (ref, dist):
                planes = []
                for p in img:
                    ptr = ctypes.c_void_p()
                    S.check(lib.Vship_PinnedMalloc(ctypes.byref(ptr), p.nbytes))
                    pinned.append(ptr)
                    ctypes.memmove(ptr, p.ctypes.data, p.nbytes)
                    planes.append(ptr.value)
                pair.append(planes)
            host_pairs.append(pair)
        host_args = [(S._plane_args(pr[0], strides), S._plane_args(pr[1], strides)) for pr in host_pairs]
        pair_bytes = 2 * sum(p.nbytes for p in frames[0][0])
        for n in handler_counts:
            handlers = [S.SSIMU2Handler(cs, cs) for _ in range(n)]

            def host_worker(k, count):
                lib_ = S.load_library()
                sc = ctypes.c_double()
                for i in range(k, count, n):
                    (p1, l1), (p2, l2) = host_args[i % len(host_args)]
                    S.check(lib_.Vship_ComputeSSIMU2(handlers[k].handle, ctypes.byref(sc), p1, p2, l1, l2))

            def run_host(count):
                ths = [threading.Thread(target=host_worker, args=(k, count)) for k in range(n)]
                for t in ths:
                    t.start()
                for t in ths:
                    t.join()

            try:
                run_host(2 * n)  # the first host-plane call of a handler allocates its upload buffers
                t0 = time.perf_counter()
                run_host(n_host)
                rate = n_host / (time.perf_counter() - t0)
            finally:
                for h in handlers:
                    h.close()
            out[str(n)] = {"pairs_per_s": round(rate, 2), "h2d_gbs": round(rate * pair_bytes / 1e9, 2)}
            log(f"pcie-inclusive Vship_ComputeSSIMU2, {n} handlers: {rate:.1f} pairs/s "
                f"({rate * pair_bytes / 1e9:.2f} GB/s host to device)")
    finally:
        for p in pinned:
            lib.Vship_PinnedFree(p)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--handlers", type=int, default=8)  # FFVship's default -g (GPU streams)
    # HIP hardware queues of this process (GPU_MAX_HW_QUEUES, read when the
    # runtime starts, before the library's first call). HIP's default of 4
    # makes the 8 handler streams share queues; the plain path is also
    # measured at 4 in a child process (--no-queue-compare skips it).
    ap.add_argument("--hw-queues", type=int, default=16)
    ap.add_argument("--no-queue-compare", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="only the plain leg (the 4-queue child run)")
    ap.add_argument("--cpu-pairs", type=int, default=16)
    ap.add_argument("--cpu-pairs-single", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-pairs", type=int, default=512)
    ap.add_argument("--no-screening", action="store_true")
    ap.add_argument("--no-pcie", action="store_true")
    ap.add_argument("--no-driver", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; frame pairs shard across GPUs with one "
                                   "handler or batch context per device and no collectives"}))
        sys.exit(2)
    if not 1 <= args.hw_queues <= 32:
        ap.error("--hw-queues must be in 1..32")

    # before this process's first HIP call: the child run at HIP's default of
    # 4 hardware queues, and the synthesis (forked workers)
    queues4 = None
    if not args.plain_only and not args.no_queue_compare and args.hw_queues != 4:
        cmd = [sys.executable, os.path.abspath(__file__), "--plain-only", "--hw-queues", "4", "--steps",
               str(args.steps), "--warmup", str(args.warmup), "--handlers", str(args.handlers)]
        log("plain leg at 4 hardware queues (child process)")
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError(f"4-queue child run failed with {r.returncode}")
        queues4 = json.loads(r.stdout.strip().splitlines()[-1])
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)

    t0 = time.perf_counter()
    frames = make_frames(PAIRS)
    log(f"generated {PAIRS} pairs in {time.perf_counter() - t0:.1f}s ({WORKERS} worker processes)")
    driver = None
    if not args.plain_only and not args.no_driver:
        driver = driver_leg(frames)
    mixed_data = None
    if not args.plain_only and not args.no_mixed:
        mixed_data = make_mixed(args.mixed_pairs)

    buf, ptrs, strides = upload(frames)
    cs = S.colorspace(W, H)
    handlers = [S.SSIMU2Handler(cs, cs) for _ in range(args.handlers)]
    scores = [0.0] * PAIRS
    call_args = plain_args(ptrs, strides)
    elapsed = run_plain_timed(handlers, call_args, scores, args.steps, args.warmup)
    value = PAIRS * args.steps / elapsed
    log(f"plain fp32 ({args.hw_queues} hardware queues): {value:.1f} pairs/s, mean score {np.mean(scores):.4f}")
    if args.plain_only:
        for h in handlers:
            h.close()
        buf.free()
        print(json.dumps({"plain_pairs_per_s": round(value, 2), "hip_hw_queues": args.hw_queues,
                          "handlers": args.handlers, "ms_per_step": round(elapsed / args.steps * 1e3, 3)}),
              flush=True)
        return

    # per-kernel device time of one pair, HIP events on the handler's stream
    h0 = handlers[0]
    h0.compute_device(ptrs[0][0], ptrs[0][1], strides, strides)
    ms_front, ms_blur, ms_score = h0.profile(50)
    bytes_pair, flops_pair = h0.work()
    # algorithmic HBM bytes per launch (one pair): the front-end reads both
    # raw images and writes both images' A computed-map pixels; the blur/map
    # kernel reads them once and writes 32 B per item
    a_px = 3491280.0
    in_bytes = sum(p.nbytes for p in frames[0][0])
    front_bytes = 2.0 * in_bytes + 2.0 * 4.0 * a_px
    blur_bytes = 2.0 * 4.0 * a_px
    traffic = {}
    traffic_file = os.path.join(ROOT, "profiles", "kernel_traffic_1080p.json")
    if os.path.exists(traffic_file):
        traffic = json.load(open(traffic_file)).get("bytes_per_launch", {})
    counters = load_counters()
    plain_wl = "plain_fp32_1080p"
    per_kernel = {}
    for name, ms, nbytes in ((FRONT_KERNEL, ms_front, front_bytes), (BLUR_KERNEL, ms_blur, blur_bytes)):
        per_kernel[name] = {
            "ms": round(ms, 4), "algorithmic_bytes": nbytes,
            "hbm_achieved_gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": traffic.get(name),
            "valu": valu_roof(counters, plain_wl, name, ms, 1.0)}
    per_kernel[BLUR_KERNEL]["fir_flops_tflops"] = round(flops_pair / (ms_blur * 1e-3) / 1e12, 2)
    per_kernel["score_host"] = {"ms": round(ms_score, 4)}  # FP64 completion from the pinned partial sums
    dominant = BLUR_KERNEL if ms_blur >= ms_front else FRONT_KERNEL
    d = per_kernel[dominant]
    v = d["valu"] or {}
    # both kernels are bound by vector-ALU issue, not by HBM: the line states
    # the VALU issue fraction (live time, counted instructions) beside HBM's
    roofline = {
        "kernel": dominant, "bound": "valu",
        "achieved": v.get("achieved_tops"), "peak": v.get("peak_tops"), "unit": "T VALU lane-ops/s",
        "frac": v.get("frac"), "traffic": d["traffic"],
        "hbm": {"achieved": d["hbm_achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["hbm_frac"],
                "algorithmic_bytes": d["algorithmic_bytes"]},
        "per_kernel": per_kernel, "path_bytes_per_pair": bytes_pair,
        "path_valu": path_valu(counters, plain_wl, (FRONT_KERNEL, BLUR_KERNEL), value),
        "counters": os.path.relpath(COUNTERS_FILE, ROOT) if counters else None,
    }

    # batched FP16 storage path on the same 240 pairs: one graph replay per step
    batched = None
    if not args.no_batched:
        cs_list = [cs] * PAIRS
        p1 = [ptrs[i][0] for i in range(PAIRS)]
        p2 = [ptrs[i][1] for i in range(PAIRS)]
        st = [strides] * PAIRS
        with S.SSIMU2Batch(PAIRS, S.STORAGE_FP16) as bt:
            tuning, tune_ms = bt.autotune(cs_list, cs_list, p1, p2, st, st, reps=2)
            log(f"batched fp16 tuning {tuning} ({tune_ms:.3f} ms per batch)")
            # the timed calls reuse one set of ctypes arrays, as a C caller would
            prepared = bt.prepare(cs_list, cs_list, p1, p2, st, st)
            bscores = np.zeros(PAIRS, dtype=np.float64)
            for _ in range(args.warmup):
                bt.compute_prepared(prepared, bscores)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                bt.compute_prepared(prepared, bscores)
            b_elapsed = time.perf_counter() - t0
            b_front, b_blur, b_score, b_graph = bt.profile(10)
            b_bytes, b_flops = bt.work()
        b_value = PAIRS * args.steps / b_elapsed
        blur_bytes16 = 2.0 * 2.0 * 3491280.0 * PAIRS
        batch_wl = "batch_fp16_1080p"
        batched = {
            "value": round(b_value, 2), "unit": "pairs/s", "ratio_vs_plain": round(b_value / value, 3),
            "storage": "fp16 planes, fp32 arithmetic, fp64 norms and score",
            "tuning": tuning, "max_abs_delta_vs_plain": round(float(np.max(np.abs(np.array(bscores) - np.array(scores)))), 6),
            "kernel_ms_per_batch": {"frontend": round(b_front, 4), "blur": round(b_blur, 4), "score": round(b_score, 4),
                                    "graph": round(b_graph, 4)},
            "blur_hbm_gbs": round(blur_bytes16 / (b_blur * 1e-3) / 1e9, 1),
            "blur_hbm_frac": round(blur_bytes16 / (b_blur * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "blur_valu": valu_roof(counters, batch_wl, "blur_strip_kernel<__half>", b_blur, PAIRS),
            "frontend_valu": valu_roof(counters, batch_wl, "frontend_kernel<__half, 1, 0>", b_front, PAIRS),
            "path_valu": path_valu(counters, batch_wl, ("frontend_kernel<__half, 1, 0>", "blur_strip_kernel<__half>"),
                                   b_value),
            "blur_fir_flops_tflops": round(b_flops / (b_blur * 1e-3) / 1e12, 2),
            "path_gbs": round(b_bytes / (b_graph * 1e-3) / 1e9, 1),
        }
        log(f"batched fp16: {b_value:.1f} pairs/s ({b_value / value:.2f}x plain)")

    for h in handlers:
        h.close()
    pcie = None if args.no_pcie else pcie_leg(frames, strides, cs)

    mixed = None
    if mixed_data is not None:
        mixed = mixed_leg(*mixed_data, args.steps, args.warmup)
        mixed_data = None

    screening = None
    if not args.no_screening:
        screening = screening_leg(max(1, min(args.steps, 3)))

    cpu = None
    if not args.no_cpu_baseline:
        nproc = host_cores()
        rate1, vals1 = cpu_baseline(frames, 1, args.cpu_pairs_single)
        rate_n, vals_n = cpu_baseline(frames, WORKERS, args.cpu_pairs)
        # each sample is pairs 0, 1, ... of the step: compare pair i with the GPU's pair i
        err = max(abs(a - scores[i % PAIRS]) for vals in (vals_n, vals1) for i, a in enumerate(vals))
        cpu = {"value": round(rate_n, 3), "unit": "pairs/s", "cores": WORKERS, "kind": "port",
               "sample": f"{args.cpu_pairs} of the 240 1080p pairs, FP64 oracle, one pair per thread, "
                         f"{WORKERS} threads ({nproc} cores visible; one GPU's share of the box is 16)",
               "nproc": nproc,
               "single_core": {"value": round(rate1, 3), "unit": "pairs/s", "cores": 1,
                               "sample": f"{args.cpu_pairs_single} of the 240 1080p pairs, one thread"},
               "max_abs_delta_vs_gpu": round(err, 6)}
        log(f"cpu baseline: {rate1:.3f} pairs/s on 1 core, {rate_n:.3f} pairs/s on {WORKERS} of {nproc} cores")

    buf.free()
    out = {
        "metric": "SSIMU2 frame-pairs/s at 1080p (batched FP16 vs plain FP32); % HBM peak",
        "value": round(value, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "240 pairs 1920x1080 YUV420 8-bit BT.709 limited, DCT-quantised q~U{4..32}, "
                               "plain FP32 path, device-resident planes", "pairs_per_step": PAIRS,
                   "handlers": args.handlers, "hip_hw_queues": args.hw_queues},
        "plain_at_4_hw_queues": queues4,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched_fp16": batched,
        "batched_mixed_fp16": mixed,
        "surrogate_screening": screening,
        "pcie_inclusive": pcie,
        "ffvship_driver": driver,
        "pcie_inclusive_pairs_per_s": pcie[str(args.handlers)]["pairs_per_s"] if pcie and str(args.handlers) in pcie
        else None,
        "mean_score": round(float(np.mean(scores)), 5),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
