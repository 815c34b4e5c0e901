#!/usr/bin/env python3
"""Throughput of the threshold -> quads front-end on one MI355X.

Workload (the headline): plain per-frame pipeline on 1920x1080 uint8 frames,
30 tags + 200 rectangles/ellipses + 100 lines, blur 1.0, noise 6, seeds
2000..2099 (100 frames, generated here and uploaded to HBM before timing),
decimate 1. A step = one frame through threshold -> quads (all stages, the
detector's buffers allocated once beforehand). K steps are timed between two
device synchronisations; value = input Mpix/s = K * 1920*1080 / wall seconds.

Also reported: decimate-2 throughput on the same frames, per-stage times from
HIP events on the launch stream, the 4096x3072 stress frame (threshold and
CCL stage rooflines), and the CPU oracle (the C++ restatement of the
reference, OpenMP over the host cores) on a bounded sample of the same frames.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "sycl-apriltag-test_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
W1080, H1080 = 1920, 1080
COUNT_KEYS = ("n_boundary", "n_valid_pairs", "n_clusters", "n_cluster_points", "n_corners",
              "n_segments", "n_quads", "sort_passes", "rank_shift")
METRIC = "Mpix/s & frames/s threshold→quads (1080p, 12MP); batched vs plain; HBM GB/s %"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(stage, f, npx, counts):
    """Algorithmic HBM bytes of one stage for one frame (npx = decimated px):
    what the stage must read and write at least, with every intermediate in
    HBM once. The cluster stage has no kernels of its own (the ranking and
    the key sort's last pass produce the clusters): None."""
    nbp = int(counts["n_boundary"])
    ncp = int(counts["n_cluster_points"])
    nc = int(counts["n_corners"])
    if stage == "threshold":
        # f input rows touched per output row, thresholded write, decimated grey write
        return f * npx + npx + (npx if f > 1 else 0)
    if stage == "ccl":
        return 5 * npx  # thresholded read + u32 labels write
    if stage == "boundary":
        return 4 * npx + 12 * nbp  # labels read + (u64 key, u32 record) written
    if stage == "key_sort":
        # per pass: key + record read and written; the last pass also writes
        # each record's cluster point (8 B)
        return 24 * nbp * sort_passes(npx, counts) + 8 * nbp
    if stage == "angle_fit":
        return 80 * ncp  # point read, sorted point (8 B) and line-fit prefix (64 B) written
    if stage == "fit_lines":
        return 76 * ncp  # line-fit prefix read, corner slot (12 B) written
    if stage == "corners":
        return 12 * ncp + 24 * nc  # corner slots read, corners compacted and sorted
    return None


def stage_traffic(stage, workload="1080p_f1"):
    """HBM bytes per frame of `stage`, from the two counter runs (FETCH_SIZE,
    WRITE_SIZE, scaled by a device copy) summarised in
    profiles/pipeline_<workload>_hbm_traffic.json; None if absent."""
    path = os.path.join(ROOT, "profiles", f"pipeline_{workload}_hbm_traffic.json")
    try:
        with open(path) as f:
            st = json.load(f)["stages"].get(stage)
    except (OSError, ValueError, KeyError):
        return None
    return None if st is None else int(st["fetch_bytes"] + st["write_bytes"])


def sort_passes(npx, counts):
    """Radix passes of the frame's boundary key sort: the detector's own
    count over the pair-rank keys, or, when it fell back to the full-key
    sort (0), 2 label fields + 2 plane bits in 11-bit digits."""
    n = int(counts["sort_passes"])
    if n:
        return n
    lb = 1
    while lb < 31 and (1 << lb) < npx:
        lb += 1
    return (2 * lb + 2 + 10) // 11


def time_frames(atq, det, devs, w, h, steps, warmup, stream):
    for i in range(warmup):
        det.detect(devs[i % len(devs)], w, h, stream.handle)
    stream.sync()
    e0, e1 = atq.Event(), atq.Event()
    t0 = time.perf_counter()
    e0.record(stream.handle)
    for i in range(steps):
        det.detect(devs[i % len(devs)], w, h, stream.handle)
    e1.record(stream.handle)
    e1.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_ms(e1) / 1e3


def frame_times(atq, det, devs, w, h, stream):
    """Milliseconds of each frame of devs run once, one HIP event pair per
    frame on the launch stream (outside the timed region)."""
    ev = [atq.Event() for _ in range(len(devs) + 1)]
    ev[0].record(stream.handle)
    for i, d in enumerate(devs):
        det.detect(d, w, h, stream.handle)
        ev[i + 1].record(stream.handle)
    ev[-1].synchronize()
    return [ev[i].elapsed_ms(ev[i + 1]) for i in range(len(devs))]


def fits_and_quads(det, counts):
    """The last frame's fitted quads and quads (for the float64 corner check
    of the CPU leg)."""
    from pyatq import records as R
    return (det.read("FITTED", R.FITTED_QUAD, int(counts["n_segments"])),
            det.read("QUADS", R.QUAD_CORNERS, int(counts["n_quads"])))


def median(xs):
    xs = sorted(xs)
    n = len(xs)
    return xs[n // 2] if n % 2 else 0.5 * (xs[n // 2 - 1] + xs[n // 2])


def stage_profile(atq, det, devs, w, h, n, stream):
    det.set_timing(True)
    acc = None
    for i in range(n):
        det.detect(devs[i % len(devs)], w, h, stream.handle)
        st = det.stage_times()
        acc = st if acc is None else {k: acc[k] + v for k, v in st.items()}
    det.set_timing(False)
    return {k: v / n for k, v in acc.items()}


def tuned_shapes(det, dev, w, h, stream, tune):
    """The autotuner's shapes for this frame's resolution (or the defaults)."""
    return det.autotune(dev, w, h, stream.handle) if tune else det.shapes()


def bench_12mp(atq, synth, stream, steps=10, tune=True):
    """The 4096x3072 stress frame (maze, dense clutter), decimate 1."""
    t = time.time()
    big = synth.stress_12mp().image
    log(f"generated 12MP frame in {time.time() - t:.1f}s")
    dbig = [atq.dev_from_np(big)]
    det = atq.Detector(4096, 3072, decimate=1)
    shapes = tuned_shapes(det, dbig[0], 4096, 3072, stream, tune)
    wall, ev = time_frames(atq, det, dbig, 4096, 3072, steps, 2, stream)
    stages = stage_profile(atq, det, dbig, 4096, 3072, 5, stream)
    det.detect(dbig[0], 4096, 3072, stream.handle)
    stream.sync()
    c = det.counts()
    npx12 = 4096 * 3072
    roof12 = {}
    for st in ("threshold", "ccl", "boundary", "key_sort", "angle_fit", "fit_lines", "corners"):
        b = stage_bytes(st, 1, npx12, c)
        roof12[st] = {"ms": round(stages[st], 4), "GB/s": round(b / (stages[st] * 1e-3) / 1e9, 1),
                      "frac": round(b / (stages[st] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "algorithmic_bytes": int(b), "traffic": stage_traffic(st, "12mp")}
    out = {
        "autotuned_shapes": list(shapes),
        "ms_per_frame": round(wall / steps * 1e3, 3),
        "mpix_per_s": round(steps * npx12 / wall / 1e6, 1),
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "stage_roofline": roof12,
        "counts": {k: int(c[k]) for k in COUNT_KEYS},
    }
    log(f"12MP: {wall / steps * 1e3:.2f} ms/frame; stages {stages}")
    fits = fits_and_quads(det, c)
    det.close()
    return out, fits


def bench_batch(atq, synth, stream, rounds):
    """16 frames of mixed sizes and decimation: one detector per frame run one
    after another on one stream (plain) vs the batch's single graph launch."""
    frames = synth.batch_mixed(4000, n=16)
    sizes = [(f.image.shape[1], f.image.shape[0]) for f in frames]
    decs = [f.decimate for f in frames]
    devs = [atq.dev_from_np(f.image) for f in frames]
    dets = [atq.Detector(w, h, decimate=d) for (w, h), d in zip(sizes, decs)]
    for (w, h), det, dv in zip(sizes, dets, devs):  # the same tuned shapes as the batch
        det.autotune(dv, w, h, stream.handle)

    def plain_round():
        for (w, h), det, dv in zip(sizes, dets, devs):
            det.detect(dv, w, h, stream.handle)

    b = atq.Batch(sizes, decs)
    for i, f in enumerate(frames):
        b.upload(i, f.image)

    def timed(fn):
        for _ in range(3):
            fn()
        stream.sync()
        t0 = time.perf_counter()
        for _ in range(rounds):
            fn()
        stream.sync()
        return time.perf_counter() - t0

    t_plain = timed(plain_round)
    t_graph = timed(lambda: b.detect(stream.handle, use_graph=True))
    t_eager = timed(lambda: b.detect(stream.handle, use_graph=False))
    same = all(b.frame(i).quads().tobytes() == dets[i].quads().tobytes() for i in range(16))
    mpix = sum(w * h for w, h in sizes) / 1e6
    out = {"frames": [f"{w}x{h}/d{d}" for (w, h), d in zip(sizes, decs)],
           "rounds": rounds,
           "plain_frames_per_s": round(16 * rounds / t_plain, 1),
           "batched_graph_frames_per_s": round(16 * rounds / t_graph, 1),
           "batched_eager_frames_per_s": round(16 * rounds / t_eager, 1),
           "batched_graph_mpix_per_s": round(mpix * rounds / t_graph, 1),
           "graph_vs_plain": round(t_plain / t_graph, 3),
           "per_frame_quads_identical": bool(same),
           "autotuned_shapes": {f"{w}x{h}/d{d}": list(b.frame(i).shapes())
                                for i, ((w, h), d) in enumerate(zip(sizes, decs))}}
    log(f"batch: {out}")
    b.close()
    for det in dets:
        det.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--cpu-frames", type=int, default=20,
                    help="frames of the CPU baseline on all allowed host cores")
    ap.add_argument("--cpu-frames-1t", type=int, default=4,
                    help="frames of the CPU baseline on one thread")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-12mp", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-f2", action="store_true", help="no decimate-2 leg")
    ap.add_argument("--skip-fixed", action="store_true", help="no fixed-mode leg")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first copy this many MiB device to device (a counter reference "
                         "for profiler runs)")
    ap.add_argument("--only-12mp", action="store_true",
                    help="only the 4096x3072 leg, --steps frames (for profiling)")
    ap.add_argument("--batch-rounds", type=int, default=20)
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default front-end shapes (no autotuner runs, e.g. under a profiler)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on exactly one GPU (frames are independent; "
              "a batch would shard by frame across GPUs). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pyatq as atq
    from pyatq import synth

    atq.init()
    if args.copy_mb:
        nb = args.copy_mb << 20
        src, dst = atq.DevBuf(nb), atq.DevBuf(nb)
        atq.hip_check(atq._hip().hipMemcpy(atq.C.c_void_p(dst.ptr), atq.C.c_void_p(src.ptr), nb, 3),
                      "hipMemcpy D2D")
        log(f"copied {args.copy_mb} MiB device to device")
        src.free()
        dst.free()
    if args.only_12mp:
        print(json.dumps({"stress_12mp": bench_12mp(atq, synth, atq.Stream(), args.steps,
                                                     not args.no_tune)[0]}))
        return
    t = time.time()
    frames = [synth.plain_1080p(2000 + i).image for i in range(args.frames)]
    log(f"generated {len(frames)} frames in {time.time() - t:.1f}s")
    devs = [atq.dev_from_np(f) for f in frames]
    stream = atq.Stream()

    res = {}
    for f in ((1,) if args.skip_f2 else (1, 2)):
        det = atq.Detector(W1080, H1080, decimate=f)
        shapes = tuned_shapes(det, devs[0], W1080, H1080, stream, not args.no_tune)
        wall, ev = time_frames(atq, det, devs, W1080, H1080, args.steps, args.warmup, stream)
        per_frame = frame_times(atq, det, devs, W1080, H1080, stream)
        stages = stage_profile(atq, det, devs, W1080, H1080, 20, stream)
        det.detect(devs[0], W1080, H1080, stream.handle)
        stream.sync()
        counts = det.counts()
        res[f] = dict(wall=wall, ev=ev, stages=stages, counts=counts, shapes=shapes,
                      per_frame=per_frame, fits=fits_and_quads(det, counts))
        log(f"decimate {f}: {args.steps} frames in {wall * 1e3:.1f} ms wall "
            f"({ev * 1e3:.1f} ms events), quads/frame {int(counts['n_quads'])}")
        log("  stages ms: " + ", ".join(f"{k} {v:.4f}" for k, v in stages.items()))
        det.close()

    steps = args.steps
    wall1 = res[1]["wall"]
    mpix = steps * W1080 * H1080 / wall1 / 1e6
    fps = steps / wall1

    # dominant stage at the headline workload and its roofline
    npx = W1080 * H1080
    st1 = res[1]["stages"]
    dom = max(st1, key=st1.get)
    roof = None
    alg = stage_bytes(dom, 1, npx, res[1]["counts"])
    traffic = stage_traffic(dom)
    if alg is not None:
        gbs = alg / (st1[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": traffic, "stage": dom,
                "algorithmic_bytes": int(alg), "avg_ms": round(st1[dom], 5)}
    else:
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": traffic, "stage": dom, "avg_ms": round(st1[dom], 5)}

    stage_roof = {}
    for st in ("threshold", "ccl", "boundary", "key_sort", "angle_fit", "fit_lines", "corners"):
        b = stage_bytes(st, 1, npx, res[1]["counts"])
        stage_roof[st] = {"ms": round(st1[st], 5), "GB/s": round(b / (st1[st] * 1e-3) / 1e9, 1),
                          "frac": round(b / (st1[st] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    pf = res[1]["per_frame"]
    extra = {
        "frames_per_s": round(fps, 2),
        "frame_ms_median": round(median(pf), 4),
        "frames_per_s_median": round(1e3 / median(pf), 1),
        "frame_ms_p90": round(sorted(pf)[int(0.9 * (len(pf) - 1))], 4),
        "per_frame_timing": f"{len(pf)} frames (seeds 2000-{2000 + len(pf) - 1}) run once each "
                            "after the timed region, one HIP event pair per frame",
        "event_ms_per_frame": round(res[1]["ev"] / steps * 1e3, 4),
        "autotuned_shapes_1080p_f1": list(res[1]["shapes"]),
        "stage_ms_1080p_f1": {k: round(v, 5) for k, v in st1.items()},
        "stage_roofline_1080p_f1": stage_roof,
        "counts_1080p_f1": {k: int(res[1]["counts"][k]) for k in COUNT_KEYS},
    }

    if 2 in res:
        extra["mpix_per_s_decimate2"] = round(steps * W1080 * H1080 / res[2]["wall"] / 1e6, 2)
        extra["frames_per_s_decimate2"] = round(steps / res[2]["wall"], 2)
        extra["stage_ms_1080p_f2"] = {k: round(v, 5) for k, v in res[2]["stages"].items()}

    if not args.skip_fixed:
        # the detector's fixed mode (reference behaviours corrected) on the
        # same frames, for comparison; not the headline
        det = atq.Detector(W1080, H1080, decimate=1, fixed=True)
        tuned_shapes(det, devs[0], W1080, H1080, stream, not args.no_tune)
        wall, _ = time_frames(atq, det, devs, W1080, H1080, args.steps, args.warmup, stream)
        stages = stage_profile(atq, det, devs, W1080, H1080, 20, stream)
        det.detect(devs[0], W1080, H1080, stream.handle)
        stream.sync()
        c = det.counts()
        det.close()
        extra["fixed_mode_1080p_f1"] = {
            "mpix_per_s": round(steps * W1080 * H1080 / wall / 1e6, 2),
            "ms_per_frame": round(wall / steps * 1e3, 4),
            "stage_ms": {k: round(v, 5) for k, v in stages.items()},
            "n_clusters": int(c["n_clusters"]), "n_quads": int(c["n_quads"])}
        log(f"fixed mode: {extra['fixed_mode_1080p_f1']['ms_per_frame']} ms/frame")

    fits = {"1920x1080 seed 2000 f=1": res[1]["fits"]}
    if 2 in res:
        fits["1920x1080 seed 2000 f=2"] = res[2]["fits"]
    small = synth.plain_small()
    hs, ws = small.image.shape
    det = atq.Detector(ws, hs, decimate=small.decimate)
    det.detect(atq.dev_from_np(small.image), ws, hs, stream.handle)
    stream.sync()
    fits["640x480 seed 1000 f=2"] = fits_and_quads(det, det.counts())
    det.close()
    if not args.skip_12mp:
        extra["stress_12mp"], fits["4096x3072 stress f=1"] = bench_12mp(atq, synth, stream,
                                                                       tune=not args.no_tune)

    if not args.skip_batch:
        extra["batched_16_mixed"] = bench_batch(atq, synth, stream, args.batch_rounds)

    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        orc = pyoracle.Oracle()
        # the cores this process may run on, within the OpenMP thread limit
        # the host sets (the GPU box's share of its CPUs)
        nproc = len(os.sched_getaffinity(0))
        nthreads = min(nproc, orc.max_threads())

        def per_frame_s(sample, threads):
            orc.detect(sample[0], 1, threads)  # warm-up
            ts = []
            for fr in sample:
                t = time.perf_counter()
                orc.detect(fr, 1, threads)
                ts.append(time.perf_counter() - t)
            return ts

        n_all = min(args.cpu_frames, len(frames))
        t_all = per_frame_s(frames[:n_all], nthreads)
        t_one = per_frame_s(frames[:min(args.cpu_frames_1t, len(frames))], 1)
        cpu = {"value": round(npx / median(t_all) / 1e6, 3), "unit": "Mpix/s",
               "cores": nthreads, "kind": "port",
               "sample": f"{n_all} of the same 1080p frames (seeds 2000-{1999 + n_all}), decimate 1, "
                         f"median per frame; C++ restatement of the reference, OpenMP over "
                         f"{nthreads} threads (nproc {nproc}), {sum(t_all):.2f}s",
               "nproc": nproc,
               "frame_ms_median": round(median(t_all) * 1e3, 2),
               "one_thread": {"value": round(npx / median(t_one) / 1e6, 3), "unit": "Mpix/s",
                              "frame_ms_median": round(median(t_one) * 1e3, 2),
                              "sample": f"{len(t_one)} frames (seeds 2000-{1999 + len(t_one)}), "
                                        f"{sum(t_one):.2f}s"}}
        log(f"cpu oracle: {cpu}")
        # the float32 quad corners (bit-exact with the reference
        # restatement) against a float64 twin of the quad stage
        err = {}
        for name, (fitted, quads) in fits.items():
            e, n = pyoracle.corner_error_fp64(fitted, quads)
            err[name] = {"quads": n, "max_abs_px": float(f"{e:.3e}")}
        extra["corner_error_vs_float64"] = err
        log(f"corner error vs float64: {err}")

    line = {
        "metric": METRIC, "value": round(mpix, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(wall1 / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"plain per-frame threshold->quads, 1920x1080 u8, 30 tags + 200 "
                               f"shapes + 100 lines, blur 1.0 noise 6, decimate 1; {steps} timed "
                               f"frames cycling through seeds 2000-{2000 + min(steps, len(devs)) - 1}"
                               f" (throughput over the K timed frames)",
                   "global_batch": 1, "frames": args.frames, "parallelism": "single-gpu"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
