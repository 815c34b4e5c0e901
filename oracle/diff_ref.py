"""Float64 reference of the differentiable line-fit / corner stage.

TEST INFRASTRUCTURE ONLY (the gradient-check harness and bench.py's
differentiation leg). The forward restates the C ABI's definition
(include/atq.h, atq_linefit_corners_forward) in float64 with two-pass
moments; the backward is the analytic vector-Jacobian product, written here
independently of the HIP kernel and checked against central finite
differences by tests/test_diff.py.

Point i of a cluster sits at (X, Y) = (x + 1, y + 0.5), weight
w_i = s_i * sqrt(1 + gx^2 + gy^2) with gx = I[y, x+2] - I[y, x],
gy = I[y+1, x+1] - I[y-1, x+1] (zero when (x+1, y) touches the border).
Line k fits the cyclic range [idx_k, idx_{k+1}]; corner k intersects lines k
and k+1 mod 4.
"""
import numpy as np


def point_arrays(points):
    x = (points["packed_x"].astype(np.int64) & 0x3FFF)
    y = (points["packed_y"].astype(np.int64) & 0x3FFF)
    return x + 1, y, (x + 1).astype(np.float64), y.astype(np.float64) + 0.5


def point_weights(img, points, s):
    """w, r, gx, gy, pixel (px, py) of every point (float64)."""
    h, w = img.shape
    px, py, _, _ = point_arrays(points)
    ok = (px > 0) & (px + 1 < w) & (py > 0) & (py + 1 < h)
    pxc, pyc = np.clip(px, 1, w - 2), np.clip(py, 1, h - 2)
    gx = np.where(ok, img[pyc, pxc + 1] - img[pyc, pxc - 1], 0.0)
    gy = np.where(ok, img[pyc + 1, pxc] - img[pyc - 1, pxc], 0.0)
    r = np.sqrt(1.0 + gx * gx + gy * gy)
    return s * r, r, gx, gy, px, py, ok


def cyclic(i0, i1, n):
    return np.arange(i0, i1 + 1) if i0 <= i1 else np.r_[np.arange(i0, n), np.arange(0, i1 + 1)]


def _line(X, Y, w, sel):
    ws, xs, ys = w[sel], X[sel], Y[sel]
    W = ws.sum()
    mx, my = (ws * xs).sum() / W, (ws * ys).sum() / W
    u, v = xs - mx, ys - my
    cxx, cyy, cxy = (ws * u * u).sum() / W, (ws * v * v).sum() / W, (ws * u * v).sum() / W
    A, B = cxx - cyy, 2.0 * cxy
    phi = 0.5 * np.arctan2(B, A)
    return dict(W=W, mx=mx, my=my, A=A, B=B, d=np.array([np.cos(phi), np.sin(phi)]), sel=sel)


def quad_forward(img, points, quad, s):
    """corners (4, 2) of one quad; points/s are the cluster's slices."""
    n = int(quad["cluster_count"])
    idx = [int(i) for i in quad["idx"]]
    w, *_ = point_weights(img, points, s)
    _, _, X, Y = point_arrays(points)
    L = [_line(X, Y, w, cyclic(idx[k], idx[(k + 1) & 3], n)) for k in range(4)]
    out = np.zeros((4, 2))
    for k in range(4):
        a, b = L[k], L[(k + 1) & 3]
        ma, mb = np.array([a["mx"], a["my"]]), np.array([b["mx"], b["my"]])
        D = a["d"][0] * b["d"][1] - a["d"][1] * b["d"][0]
        t = ((mb - ma)[0] * b["d"][1] - (mb - ma)[1] * b["d"][0]) / D
        out[k] = ma + t * a["d"]
    return out, L


def quad_backward(img, points, quad, s, g):
    """dL/ds (cluster slice) and dL/dI (image) for upstream g = dL/dcorners (4, 2)."""
    n = int(quad["cluster_count"])
    corners, L = quad_forward(img, points, quad, s)
    w, r, gx, gy, px, py, ok = point_weights(img, points, s)
    _, _, X, Y = point_arrays(points)
    gm = np.zeros((4, 2))
    gph = np.zeros(4)
    for k in range(4):
        a, b = L[k], L[(k + 1) & 3]
        ia, ib = k, (k + 1) & 3
        ma, mb = np.array([a["mx"], a["my"]]), np.array([b["mx"], b["my"]])
        da, db = a["d"], b["d"]
        D = da[0] * db[1] - da[1] * db[0]
        t = ((mb - ma)[0] * db[1] - (mb - ma)[1] * db[0]) / D
        gc = g[k]
        G1 = gc @ da
        dadb = da @ db
        gm[ia] += gc + G1 * np.array([-db[1], db[0]]) / D
        gm[ib] += G1 * np.array([db[1], -db[0]]) / D
        na = np.array([-da[1], da[0]])
        gph[ia] += t * (gc @ na) + G1 * t * dadb / D
        gph[ib] += G1 * (((mb - ma) @ db) - t * dadb) / D
    dLdw = np.zeros(n)
    for k in range(4):
        Lk = L[k]
        sel = Lk["sel"]
        u, v = X[sel] - Lk["mx"], Y[sel] - Lk["my"]
        dA = (u * u - v * v - Lk["A"]) / Lk["W"]
        dB = (2 * u * v - Lk["B"]) / Lk["W"]
        dphi = 0.5 * (Lk["A"] * dB - Lk["B"] * dA) / (Lk["A"] ** 2 + Lk["B"] ** 2)
        np.add.at(dLdw, sel, gm[k, 0] * u / Lk["W"] + gm[k, 1] * v / Lk["W"] + gph[k] * dphi)
    gs = dLdw * r
    gimg = np.zeros_like(img)
    dx = np.where(ok, dLdw * s * gx / r, 0.0)
    dy = np.where(ok, dLdw * s * gy / r, 0.0)
    np.add.at(gimg, (py, px + 1), dx)
    np.add.at(gimg, (py, px - 1), -dx)
    np.add.at(gimg, (py + 1, px), dy)
    np.add.at(gimg, (py - 1, px), -dy)
    return gs, gimg


def forward_all(img, points, quads, s):
    out = np.zeros((len(quads), 4, 2))
    for q, Q in enumerate(quads):
        a, n = int(Q["cluster_start"]), int(Q["cluster_count"])
        out[q], _ = quad_forward(img, points[a:a + n], Q, s[a:a + n])
    return out


def backward_all(img, points, quads, s, g):
    gs = np.zeros(len(s))
    gimg = np.zeros_like(img)
    for q, Q in enumerate(quads):
        a, n = int(Q["cluster_start"]), int(Q["cluster_count"])
        dgs, dgi = quad_backward(img, points[a:a + n], Q, s[a:a + n], g[q])
        gs[a:a + n] += dgs
        gimg += dgi
    return gs, gimg
