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


EDGE_T = np.linspace(0.15, 0.85, 16)


def bilinear(img, X, Y):
    """Bilinear image value and its (X, Y) gradient; pixel (x, y) is the
    square [x, x+1) x [y, y+1), so its value sits at (x + 0.5, y + 0.5)."""
    h, w = img.shape
    u, v = X - 0.5, Y - 0.5
    x0 = np.clip(np.floor(u).astype(np.int64), 0, w - 2)
    y0 = np.clip(np.floor(v).astype(np.int64), 0, h - 2)
    fx, fy = u - x0, v - y0
    i00, i10, i01, i11 = img[y0, x0], img[y0, x0 + 1], img[y0 + 1, x0], img[y0 + 1, x0 + 1]
    val = i00 * (1 - fx) * (1 - fy) + i10 * fx * (1 - fy) + i01 * (1 - fx) * fy + i11 * fx * fy
    dx = (i10 - i00) * (1 - fy) + (i11 - i01) * fy
    dy = (i01 - i00) * (1 - fx) + (i11 - i10) * fx
    return val, dx, dy


def _edge_samples(c):
    """Points along the four edges of each quad (n, 4, len(EDGE_T), 2); edge k
    runs from corner k-1 to corner k."""
    a = np.roll(c, 1, axis=1)
    t = EDGE_T[None, None, :, None]
    return a[:, :, None, :] * (1 - t) + c[:, :, None, :] * t, a


def edge_mid_level(img, c):
    """Per quad the grey level halfway between the two sides of its edges,
    from samples 1.5 px either side of each edge of the corners c (n, 4, 2)."""
    p, a = _edge_samples(c)
    d = c - a
    n = np.stack([-d[..., 1], d[..., 0]], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    off = 1.5 * n[:, :, None, :]
    hi = bilinear(img, (p + off)[..., 0], (p + off)[..., 1])[0]
    lo = bilinear(img, (p - off)[..., 0], (p - off)[..., 1])[0]
    return 0.5 * (hi.mean(axis=(1, 2)) + lo.mean(axis=(1, 2)))


def edge_straddle_loss(img, c, mid):
    """Loss that needs no ground truth: the squared difference between the
    image along each fitted edge and the quad's mid level (an edge that sits
    on the intensity step reads the mid level everywhere). Returns the loss
    and its gradient with respect to the corners (n, 4, 2)."""
    p, _ = _edge_samples(c)
    val, dx, dy = bilinear(img, p[..., 0], p[..., 1])
    r = val - mid[:, None, None]
    gp = np.stack([2 * r * dx, 2 * r * dy], -1)
    t = EDGE_T[None, None, :, None]
    g_end = (gp * t).sum(2)
    g_start = (gp * (1 - t)).sum(2)
    return float((r ** 2).sum()), g_end + np.roll(g_start, -1, axis=1)


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
