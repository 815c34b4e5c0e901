"""Compat-versus-fixed report: the detector's two modes on synthetic frames
whose tag corners are known.

For every frame and mode: the valid clusters and quads the detector finds,
how many of the frame's tags a quad matches (mean corner distance below
2 px in the input frame, for each of the quad's 8 corner orderings), the
corner error of those matches, and the GPU time per frame. Corners of a
decimated frame map back as upstream apriltag maps them: (c - 0.5) * f + 0.5.

Usage: python -m pyatq.modes [--reps N] [--quick]
Prints a markdown table.
"""
from __future__ import annotations

import argparse
import time

import numpy as np

MATCH_PX = 2.0


def corner_errors(quads, truth, f=1):
    """Per ground-truth tag, the smallest mean corner distance (input pixels)
    over the quads and their 8 corner orderings; inf without quads."""
    out = np.full(len(truth), np.inf)
    if len(quads) == 0 or len(truth) == 0:
        return out
    q = quads["corners"].astype(np.float64)
    if f > 1:
        q = (q - 0.5) * f + 0.5
    orders = []
    for r in range(4):
        orders.append(np.roll(np.arange(4), r))
        orders.append(np.roll(np.arange(4)[::-1], r))
    cand = np.stack([q[:, o] for o in orders], 1)  # (n_quads, 8, 4, 2)
    for i, t in enumerate(truth):
        d = np.sqrt(((cand - np.asarray(t, np.float64)) ** 2).sum(-1)).mean(-1)
        out[i] = d.min()
    return out


def frames(quick=False):
    """(name, image, decimate, truth corners) of the report's workloads."""
    from . import synth
    out = []
    fr = synth.diff_frame()
    out.append(("1280x720 supersampled (24 tags)", fr.image, 1, fr.corners))
    seeds = (2000, 2001) if quick else (2000, 2001, 2002, 2003)
    for f in (1, 2):
        for s in seeds:
            fr = synth.plain_1080p(s)
            out.append((f"1920x1080 seed {s}, f={f} ({len(fr.corners)} tags)", fr.image, f,
                        fr.corners))
    if not quick:
        fr = synth.stress_12mp()
        out.append((f"4096x3072 stress ({len(fr.corners)} tags)", fr.image, 1, fr.corners))
    return out


def run_mode(atq, img, f, fixed, reps):
    h, w = img.shape
    det = atq.Detector(w, h, decimate=f, fixed=fixed)
    dev = atq.dev_from_np(img)
    det.detect(dev, w, h)
    atq.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        det.detect(dev, w, h)
    atq.sync()
    ms = (time.perf_counter() - t0) * 1e3 / max(reps, 1)
    c = det.counts()
    quads = det.quads()
    det.close()
    return {"clusters": int(c["n_clusters"]), "quads": quads, "ms": ms}


def report(reps=20, quick=False):
    import pyatq as atq
    atq.init()
    rows = []
    for name, img, f, truth in frames(quick):
        for fixed in (False, True):
            r = run_mode(atq, img, f, fixed, reps)
            err = corner_errors(r["quads"], truth, f)
            hit = err[err < MATCH_PX]
            rows.append({
                "frame": name, "mode": "fixed" if fixed else "compat",
                "clusters": r["clusters"], "quads": len(r["quads"]),
                "found": int(len(hit)), "tags": len(truth),
                "mean": float(hit.mean()) if len(hit) else float("nan"),
                "median": float(np.median(hit)) if len(hit) else float("nan"),
                "p90": float(np.percentile(hit, 90)) if len(hit) else float("nan"),
                "ms": r["ms"],
            })
    return rows


def markdown(rows):
    lines = ["| frame | mode | valid clusters | quads | tags found | mean corner error (px) "
             "| median | p90 | ms/frame |",
             "|---|---|---:|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        lines.append(f"| {r['frame']} | {r['mode']} | {r['clusters']} | {r['quads']} | "
                     f"{r['found']}/{r['tags']} | {r['mean']:.3f} | {r['median']:.3f} | "
                     f"{r['p90']:.3f} | {r['ms']:.3f} |")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--reps", type=int, default=20, help="timed detections per frame and mode")
    ap.add_argument("--quick", action="store_true", help="fewer frames, no 12 MP frame")
    a = ap.parse_args(argv)
    print(markdown(report(a.reps, a.quick)))


if __name__ == "__main__":
    main()
