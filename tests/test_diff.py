"""Differentiable line-fit / corner stage (the differentiation form).

CPU: the float64 analytic vector-Jacobian product (oracle/diff_ref.py)
against central finite differences (float64, step 1e-3) on the boundary
weights and on the pixel intensities.
GPU, on 256 candidates from the 1280x720 frames: the HIP forward against
the float64 forward, the HIP backward against the float64 analytic gradient
(relative error <= 1e-4), the same through the torch.autograd.Function
wrapper, and 10 gradient steps on the log boundary weights that lower a loss
computed from the image alone (the ground truth only scores the result).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import diff_ref  # noqa: E402
from diff_harness import (collect_candidates, edge_mid_level,  # noqa: E402
                          edge_straddle_loss)

FD_STEP = 1e-3


@pytest.fixture(scope="module")
def cands(oracle):
    return collect_candidates(oracle, n_wanted=40, max_frames=4)


@pytest.fixture(scope="module")
def cands256(oracle):
    fr = collect_candidates(oracle, n_wanted=256, max_frames=60)
    assert sum(len(f["quads"]) for f in fr) == 256
    return fr


def _scalar(img, pts, quads, s, g):
    return float((diff_ref.forward_all(img, pts, quads, s) * g).sum())


def test_candidates_are_close_to_ground_truth(cands):
    err = []
    for fr in cands:
        c = diff_ref.forward_all(fr["img"], fr["points"], fr["quads"], np.ones(len(fr["points"])))
        err.append(np.abs(c - fr["gt"]).reshape(-1))
    err = np.concatenate(err)
    assert len(err) >= 8 * 20
    assert np.sqrt((err ** 2).mean()) < 1.0  # the fit is within a pixel of the truth


def test_analytic_gradient_matches_finite_differences(cands):
    rng = np.random.default_rng(0)
    fr = cands[0]
    img, pts, quads = fr["img"], fr["points"], fr["quads"][:6]
    s = rng.uniform(0.5, 1.5, len(pts))
    g = rng.normal(size=(len(quads), 4, 2))
    gs, gi = diff_ref.backward_all(img, pts, quads, s, g)
    # directional derivatives along random directions
    for _ in range(3):
        v = rng.normal(size=len(pts))
        fd = (_scalar(img, pts, quads, s + FD_STEP * v, g) -
              _scalar(img, pts, quads, s - FD_STEP * v, g)) / (2 * FD_STEP)
        an = gs @ v
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)
        vi = np.zeros_like(img)
        mask = gi != 0
        vi[mask] = rng.normal(size=mask.sum())
        fd = (_scalar(img + FD_STEP * vi, pts, quads, s, g) -
              _scalar(img - FD_STEP * vi, pts, quads, s, g)) / (2 * FD_STEP)
        an = (gi * vi).sum()
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)
    # element-wise: the largest weight gradients
    for i in np.argsort(-np.abs(gs))[:5]:
        e = np.zeros(len(pts))
        e[i] = FD_STEP
        fd = (_scalar(img, pts, quads, s + e, g) - _scalar(img, pts, quads, s - e, g)) / (2 * FD_STEP)
        assert abs(fd - gs[i]) <= 1e-5 * max(1e-3, abs(gs[i]))


def _gpu_forward_backward(atq, fr, s, g):
    img = fr["img"].astype(np.float32)
    h, w = img.shape
    pts, quads = fr["points"], fr["quads"]
    img_d, pts_d, q_d = atq.dev_from_np(img), atq.dev_from_np(pts), atq.dev_from_np(quads)
    s_d = atq.dev_from_np(s.astype(np.float32))
    c_d = atq.dev_empty(len(quads) * 32)
    atq.check(atq.lib().atq_linefit_corners_forward(atq.stream(), atq.P(img_d), w, h, atq.P(pts_d),
                                                    len(pts), atq.P(s_d), atq.P(q_d), len(quads),
                                                    atq.P(c_d), None))
    g_d = atq.dev_from_np(g.astype(np.float32).reshape(-1))
    gw_d = atq.dev_from_np(np.zeros(len(pts), np.float32))
    gi_d = atq.dev_from_np(np.zeros(w * h, np.float32))
    atq.check(atq.lib().atq_linefit_corners_backward(atq.stream(), atq.P(img_d), w, h, atq.P(pts_d),
                                                     len(pts), atq.P(s_d), atq.P(q_d), len(quads),
                                                     atq.P(g_d), atq.P(gw_d), atq.P(gi_d)))
    c = c_d.download(np.float32, len(quads) * 8).reshape(-1, 4, 2)
    return c, gw_d.download(np.float32, len(pts)), gi_d.download(np.float32, w * h).reshape(h, w)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
def test_gpu_forward_and_backward_match_float64(oracle, cands256):
    import pyatq as atq
    atq.init()
    rng = np.random.default_rng(1)
    worst = {"corner": 0.0, "gs": 0.0, "gi": 0.0}
    for fr in cands256:
        s = rng.uniform(0.5, 1.5, len(fr["points"]))
        g = rng.normal(size=(len(fr["quads"]), 4, 2))
        c, gs, gi = _gpu_forward_backward(atq, fr, s, g)
        c64 = diff_ref.forward_all(fr["img"], fr["points"], fr["quads"], s)
        gs64, gi64 = diff_ref.backward_all(fr["img"], fr["points"], fr["quads"], s, g)
        worst["corner"] = max(worst["corner"], float(np.abs(c - c64).max()))
        worst["gs"] = max(worst["gs"], _rel(gs, gs64))
        worst["gi"] = max(worst["gi"], _rel(gi, gi64))
    print("float32 vs float64:", worst)
    assert worst["corner"] <= 1e-3
    assert worst["gs"] <= 1e-4 and worst["gi"] <= 1e-4


@pytest.mark.gpu
def test_torch_autograd_wrapper(oracle, cands256):
    """Runs in the same process as the other GPU tests: torch maps its own HIP
    runtime beside the one libatq.so uses, and only CPU tensors are touched
    here, so torch's device runtime must stay uninitialised."""
    torch = pytest.importorskip("torch")
    import pyatq as atq
    from diff_harness import torch_function
    atq.init()
    F = torch_function()
    for fr in cands256[:4]:
        img = torch.tensor(fr["img"], requires_grad=True)
        s = torch.ones(len(fr["points"]), dtype=torch.float64, requires_grad=True)
        corners = F.apply(img, s, fr["points"], fr["quads"])
        c64 = diff_ref.forward_all(fr["img"], fr["points"], fr["quads"], np.ones(len(fr["points"])))
        assert np.abs(corners.detach().numpy() - c64).max() <= 1e-3
        mid = torch.tensor(edge_mid_level(fr["img"], c64))
        w = torch.linspace(-1.0, 1.0, 8, dtype=torch.float64).reshape(4, 2)
        loss = (((corners - torch.tensor(c64)) * w) ** 2).sum() + (corners.sum(dim=(1, 2)) * mid).sum()
        loss.backward()
        g = (2 * (corners.detach().numpy() - c64) * w.numpy() ** 2 +
             mid.numpy()[:, None, None] * np.ones_like(c64))
        gs64, gi64 = diff_ref.backward_all(fr["img"], fr["points"], fr["quads"],
                                           np.ones(len(fr["points"])), g)
        assert _rel(s.grad.numpy(), gs64) <= 1e-4
        assert _rel(img.grad.numpy(), gi64) <= 1e-4
    assert not torch.cuda.is_initialized()


def refine(atq, fr, steps=10):
    """Gradient descent on log boundary weights against edge_straddle_loss,
    which reads only the image: the fitted edges are pulled onto the
    intensity step. Returns the loss and the corner RMS error against the
    ground truth after each step."""
    theta = np.zeros(len(fr["points"]))
    zero = np.zeros((len(fr["quads"]), 4, 2))
    c0, _, _ = _gpu_forward_backward(atq, fr, np.ones(len(theta)), zero)
    mid = edge_mid_level(fr["img"], c0.astype(np.float64))
    loss, rms = [], []
    for k in range(steps + 1):
        s = np.exp(theta)
        c, _, _ = _gpu_forward_backward(atq, fr, s, zero)
        c = c.astype(np.float64)
        L, gc = edge_straddle_loss(fr["img"], c, mid)
        loss.append(L)
        rms.append(float(np.sqrt(((c - fr["gt"]) ** 2).mean())))
        if k == steps:
            break
        _, gs, _ = _gpu_forward_backward(atq, fr, s, gc)
        gtheta = gs * s
        theta -= 0.5 / max(np.abs(gtheta).max(), 1e-12) * gtheta
    return loss, rms


@pytest.mark.gpu
def test_refinement_reduces_corner_error(oracle, cands256):
    import pyatq as atq
    atq.init()
    l0, l1, r0, r1 = [], [], [], []
    for fr in cands256:
        loss, rms = refine(atq, fr)
        l0.append(loss[0])
        l1.append(loss[-1])
        r0.append(rms[0])
        r1.append(rms[-1])
    n = sum(len(f["quads"]) for f in cands256)
    print(f"{n} quads: edge loss {sum(l0):.4g} -> {sum(l1):.4g}; corner RMS vs truth "
          f"{np.mean(r0):.4f} -> {np.mean(r1):.4f} px after 10 steps")
    assert sum(l1) < sum(l0)
    assert np.mean(r1) < 0.85 * np.mean(r0)
