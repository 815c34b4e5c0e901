"""Gradient-check harness of the differentiable line-fit / corner stage.

Builds quad candidates from the config-5 frames (1280x720, 8x8-supersampled
tags with known corners) with the oracle's pipeline, matches them to the
ground truth, and wraps the C ABI forward/backward in a torch.autograd.Function
(CPU tensors; the device side goes through pyatq's HIP buffers).
"""
import numpy as np

from pyatq import records as R


def diff_quads_from_pipeline(p):
    """atq_diff_quad records from an oracle pipeline dict (fitted quads whose
    four indices lie inside the cluster they are paired with)."""
    out = []
    ext = p["extents"]
    for e, fq in enumerate(p["fitted"]):
        idx = fq["indices"].astype(np.int64)
        if not idx.any():
            continue
        start, count = int(ext[e]["start"]), int(ext[e]["count"])
        if idx.max() >= count or len(set(idx.tolist())) != 4:
            continue
        q = np.zeros(1, R.DIFF_QUAD)[0]
        q["cluster_start"], q["cluster_count"], q["idx"] = start, count, np.sort(idx)
        out.append(q)
    return np.array(out, R.DIFF_QUAD) if out else np.zeros(0, R.DIFF_QUAD)


def match_ground_truth(corners, gt_tags, max_dist=3.0):
    """For each quad (4x2) the ground-truth corner set (4x2, cyclically
    aligned, either orientation) or None when no tag is within max_dist."""
    out = []
    for c in corners:
        best, bd = None, max_dist
        for g in gt_tags:
            for cand in (g, g[::-1]):
                for k in range(4):
                    r = np.roll(cand, k, axis=0)
                    d = np.abs(r - c).max()
                    if d < bd:
                        best, bd = r, d
        out.append(best)
    return out


def collect_candidates(oracle, n_wanted=256, seed0=5000, max_frames=40):
    """(image float64, cluster points, quads, gt per quad) per frame until
    n_wanted matched candidates are gathered."""
    from pyatq import synth
    import sys
    sys.path.insert(0, __import__("os").path.join(__import__("os").path.dirname(__file__), "..", "oracle"))
    import diff_ref
    frames = []
    total = 0
    for f in range(max_frames):
        fr = synth.diff_frame(seed0 + f)
        p = oracle.pipeline(fr.image, 1)
        quads = diff_quads_from_pipeline(p)
        if len(quads) == 0:
            continue
        img = fr.image.astype(np.float64)
        pts = p["cluster_points"]
        s = np.ones(len(pts))
        c = diff_ref.forward_all(img, pts, quads, s)
        # both frames put pixel (x, y) on [x, x+1) x [y, y+1): the renderer
        # samples at pixel centres, the line fit puts a record at (x+1, y+0.5)
        gt = list(fr.corners)
        m = match_ground_truth(c, gt)
        keep = [i for i, g in enumerate(m) if g is not None]
        if not keep:
            continue
        keep = keep[: n_wanted - total]
        frames.append(dict(img=img, points=pts, quads=quads[keep],
                           gt=np.array([m[i] for i in keep]), seed=seed0 + f))
        total += len(keep)
        if total >= n_wanted:
            break
    return frames


def torch_function():
    """torch.autograd.Function over the C ABI (forward FP32 on the GPU)."""
    import torch
    import pyatq

    class LineFitCorners(torch.autograd.Function):
        @staticmethod
        def forward(ctx, image, weights, points_np, quads_np):
            h, w = image.shape
            img_d = pyatq.dev_from_np(image.detach().numpy().astype(np.float32))
            pts_d = pyatq.dev_from_np(points_np)
            q_d = pyatq.dev_from_np(quads_np)
            s_d = pyatq.dev_from_np(weights.detach().numpy().astype(np.float32))
            c_d = pyatq.dev_empty(len(quads_np) * 32)
            pyatq.check(pyatq.lib().atq_linefit_corners_forward(
                pyatq.stream(), pyatq.P(img_d), w, h, pyatq.P(pts_d), len(points_np), pyatq.P(s_d),
                pyatq.P(q_d), len(quads_np), pyatq.P(c_d), None))
            ctx.save_for_backward(image, weights)
            ctx.bufs = (img_d, pts_d, q_d, s_d, w, h, len(points_np), len(quads_np))
            out = c_d.download(np.float32, len(quads_np) * 8).reshape(len(quads_np), 4, 2)
            return torch.from_numpy(out.astype(np.float64))

        @staticmethod
        def backward(ctx, grad):
            img_d, pts_d, q_d, s_d, w, h, npts, nq = ctx.bufs
            g_d = pyatq.dev_from_np(grad.detach().numpy().astype(np.float32).reshape(-1))
            gw_d = pyatq.dev_from_np(np.zeros(npts, np.float32))
            gi_d = pyatq.dev_from_np(np.zeros(w * h, np.float32))
            pyatq.check(pyatq.lib().atq_linefit_corners_backward(
                pyatq.stream(), pyatq.P(img_d), w, h, pyatq.P(pts_d), npts, pyatq.P(s_d),
                pyatq.P(q_d), nq, pyatq.P(g_d), pyatq.P(gw_d), pyatq.P(gi_d)))
            gi = torch.from_numpy(gi_d.download(np.float32, w * h).reshape(h, w).astype(np.float64))
            gw = torch.from_numpy(gw_d.download(np.float32, npts).astype(np.float64))
            return gi, gw, None, None

    return LineFitCorners
