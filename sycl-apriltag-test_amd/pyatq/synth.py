"""Seeded synthetic grey frames for the quad front-end workloads.

Every frame is built with numpy's PCG64 generator from an integer seed, so
any workload can be regenerated bit for bit on any machine:

* ``plain_small``   640x480, 8 tags, uniform-noise background (decimate 2)
* ``plain_1080p``   1920x1080, 30 tags plus rectangle/ellipse/line clutter
* ``stress_12mp``   4096x3072, 200 tags, 5000 blobs and a 1024x1024 maze of
                    1-px corridors (worst case for union-find depth)
* ``batch_mixed``   16 frames of uneven sizes with per-frame decimation
* ``diff_frames``   1280x720 anti-aliased frames with ground-truth corners

A tag is a 10x10 cell grid: a white outer ring, the 8x8 black-bordered
square (its black border ring and a random 6x6 payload), warped into the
image by a pinhole homography with a random in-plane rotation and a
perspective tilt. ``side`` is the size of the 8x8 square in pixels.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

try:
    from scipy.ndimage import gaussian_filter as _gauss
except Exception:  # pragma: no cover - scipy is present in the image
    _gauss = None


@dataclass
class Frame:
    image: np.ndarray            # uint8 HxW
    seed: int
    decimate: int = 1
    corners: list = field(default_factory=list)  # ground truth, per tag 4x2
    payloads: list = field(default_factory=list)  # per tag its 6x6 cells (row = y), white = 1


def _homography(rng, side, cx, cy, tilt_max_deg, focal=1200.0):
    """Pinhole homography mapping tag-plane (X, Y, 1) in pixels to the image."""
    rot = rng.uniform(0.0, 2.0 * math.pi)
    tilt = math.radians(rng.uniform(0.0, tilt_max_deg))
    axis = rng.uniform(0.0, 2.0 * math.pi)
    ax, ay = math.cos(axis), math.sin(axis)
    # Rodrigues rotation about the in-plane axis (ax, ay, 0) by `tilt`.
    k = np.array([[0.0, 0.0, ay], [0.0, 0.0, -ax], [-ay, ax, 0.0]])
    r_tilt = np.eye(3) + math.sin(tilt) * k + (1.0 - math.cos(tilt)) * (k @ k)
    c, s = math.cos(rot), math.sin(rot)
    r_rot = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    r = r_rot @ r_tilt
    kmat = np.array([[focal, 0.0, cx], [0.0, focal, cy], [0.0, 0.0, 1.0]])
    rt = np.stack([r[:, 0], r[:, 1], np.array([0.0, 0.0, focal])], axis=1)
    return kmat @ rt


def _apply_h(h, pts):
    p = np.c_[pts, np.ones(len(pts))] @ h.T
    return p[:, :2] / p[:, 2:3]


def _tag_cells(rng):
    cells = np.zeros((10, 10), dtype=np.float32)
    cells[0, :] = cells[9, :] = cells[:, 0] = cells[:, 9] = 1.0  # white ring
    cells[2:8, 2:8] = rng.integers(0, 2, size=(6, 6)).astype(np.float32)
    return cells


def _render_tag(img, rng, side, tilt_max_deg, black, white, supersample=1, margin=2):
    """Draws one tag into img (float32) in place; returns its 4 image corners
    (the outer corners of the black square, corner 0 at the payload's top-left
    cell, then along the payload's rows) and its 6x6 payload, or None when it
    does not fit."""
    hgt, wid = img.shape
    half = side * 5.0 / 8.0  # half-size of the 10-cell grid
    reach = half * 1.5 + margin
    if wid <= 2 * reach or hgt <= 2 * reach:
        return None
    cx = rng.uniform(reach, wid - reach)
    cy = rng.uniform(reach, hgt - reach)
    h = _homography(rng, side, cx, cy, tilt_max_deg)
    grid = np.array([[-half, -half], [half, -half], [half, half], [-half, half]])
    quad = _apply_h(h, grid)
    x0 = max(int(math.floor(quad[:, 0].min())) - 1, 0)
    x1 = min(int(math.ceil(quad[:, 0].max())) + 2, wid)
    y0 = max(int(math.floor(quad[:, 1].min())) - 1, 0)
    y1 = min(int(math.ceil(quad[:, 1].max())) + 2, hgt)
    if x1 <= x0 or y1 <= y0:
        return None
    hinv = np.linalg.inv(h)
    cells = _tag_cells(rng)
    cell = side / 8.0
    ss = supersample
    offs = (np.arange(ss) + 0.5) / ss
    acc = np.zeros((y1 - y0, x1 - x0), dtype=np.float32)
    cov = np.zeros_like(acc)
    yy, xx = np.mgrid[y0:y1, x0:x1].astype(np.float64)
    for oy in offs:
        for ox in offs:
            px = xx + ox
            py = yy + oy
            den = hinv[2, 0] * px + hinv[2, 1] * py + hinv[2, 2]
            u = (hinv[0, 0] * px + hinv[0, 1] * py + hinv[0, 2]) / den
            v = (hinv[1, 0] * px + hinv[1, 1] * py + hinv[1, 2]) / den
            ci = np.floor((u + half) / cell).astype(np.int64)
            cj = np.floor((v + half) / cell).astype(np.int64)
            inside = (ci >= 0) & (ci < 10) & (cj >= 0) & (cj < 10)
            val = np.where(inside, cells[np.clip(cj, 0, 9), np.clip(ci, 0, 9)], 0.0)
            acc += np.where(inside, black + (white - black) * val, 0.0).astype(np.float32)
            cov += inside.astype(np.float32)
    cov /= ss * ss
    acc /= ss * ss
    region = img[y0:y1, x0:x1]
    region[:] = region * (1.0 - cov) + acc
    sq = side / 2.0
    return (_apply_h(h, np.array([[-sq, -sq], [sq, -sq], [sq, sq], [-sq, sq]])),
            cells[2:8, 2:8].astype(np.uint8))


def _render_clutter(img, rng, n_shapes, n_lines, size_lo=5, size_hi=150):
    hgt, wid = img.shape
    for _ in range(n_shapes):
        sw = rng.uniform(size_lo, size_hi)
        sh = rng.uniform(size_lo, size_hi)
        cx = rng.uniform(0, wid)
        cy = rng.uniform(0, hgt)
        val = rng.uniform(0, 255)
        ang = rng.uniform(0, math.pi)
        ellipse = rng.random() < 0.5
        r = 0.5 * math.hypot(sw, sh) + 1
        x0, x1 = max(int(cx - r), 0), min(int(cx + r) + 1, wid)
        y0, y1 = max(int(cy - r), 0), min(int(cy + r) + 1, hgt)
        if x1 <= x0 or y1 <= y0:
            continue
        yy, xx = np.mgrid[y0:y1, x0:x1].astype(np.float32)
        dx, dy = xx + 0.5 - cx, yy + 0.5 - cy
        ca, sa = math.cos(ang), math.sin(ang)
        u = ca * dx + sa * dy
        v = -sa * dx + ca * dy
        if ellipse:
            m = (u / (sw / 2)) ** 2 + (v / (sh / 2)) ** 2 <= 1.0
        else:
            m = (np.abs(u) <= sw / 2) & (np.abs(v) <= sh / 2)
        img[y0:y1, x0:x1][m] = val
    for _ in range(n_lines):
        x0f, y0f = rng.uniform(0, wid), rng.uniform(0, hgt)
        length = rng.uniform(size_lo, size_hi * 2)
        ang = rng.uniform(0, 2 * math.pi)
        x1f, y1f = x0f + length * math.cos(ang), y0f + length * math.sin(ang)
        lw = rng.uniform(1, 4)
        val = rng.uniform(0, 255)
        xa, xb = max(int(min(x0f, x1f) - lw), 0), min(int(max(x0f, x1f) + lw) + 1, wid)
        ya, yb = max(int(min(y0f, y1f) - lw), 0), min(int(max(y0f, y1f) + lw) + 1, hgt)
        if xb <= xa or yb <= ya:
            continue
        yy, xx = np.mgrid[ya:yb, xa:xb].astype(np.float32)
        px, py = xx + 0.5 - x0f, yy + 0.5 - y0f
        ddx, ddy = x1f - x0f, y1f - y0f
        t = np.clip((px * ddx + py * ddy) / (ddx * ddx + ddy * ddy), 0.0, 1.0)
        dist = np.hypot(px - t * ddx, py - t * ddy)
        img[ya:yb, xa:xb][dist <= lw / 2] = val


def _maze(rng, cells=512, corridor=220.0, wall=30.0):
    """Randomised depth-first maze of 1-px corridors, (2*cells)^2 pixels."""
    n = cells
    out = np.full((2 * n, 2 * n), wall, dtype=np.float32)
    visited = np.zeros((n, n), dtype=bool)
    dirs = ((0, 1), (1, 0), (0, -1), (-1, 0))
    stack = [(int(rng.integers(n)), int(rng.integers(n)))]
    visited[stack[0]] = True
    out[2 * stack[0][0] + 1, 2 * stack[0][1] + 1] = corridor
    order = rng.integers(0, 24, size=n * n * 2)
    perms = [p for p in __import__("itertools").permutations(range(4))]
    k = 0
    while stack:
        i, j = stack[-1]
        moved = False
        perm = perms[order[k % len(order)]]
        k += 1
        for d in perm:
            di, dj = dirs[d]
            ni, nj = i + di, j + dj
            if 0 <= ni < n and 0 <= nj < n and not visited[ni, nj]:
                visited[ni, nj] = True
                out[2 * i + 1 + di, 2 * j + 1 + dj] = corridor
                out[2 * ni + 1, 2 * nj + 1] = corridor
                stack.append((ni, nj))
                moved = True
                break
        if not moved:
            stack.pop()
    return out


def _finish(img, rng, blur, noise):
    if blur > 0:
        if _gauss is None:
            raise RuntimeError("scipy is required for the blur step")
        img = _gauss(img, blur, mode="nearest")
    if noise > 0:
        img = img + rng.normal(0.0, noise, size=img.shape).astype(np.float32)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def make_frame(seed, width, height, n_tags, side=(40, 120), tilt=30.0, bg=(80, 180),
               n_shapes=0, n_lines=0, n_blobs=0, maze=False, blur=0.8, noise=4.0,
               supersample=1, decimate=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    img = rng.integers(bg[0], bg[1] + 1, size=(height, width)).astype(np.float32)
    _render_clutter(img, rng, n_shapes, n_lines)
    if n_blobs:
        _render_clutter(img, rng, n_blobs, 0)
    if maze:
        mz = _maze(rng)
        mh, mw = mz.shape
        oy = int(rng.integers(0, height - mh + 1))
        ox = int(rng.integers(0, width - mw + 1))
        img[oy:oy + mh, ox:ox + mw] = mz
    corners, payloads = [], []
    for _ in range(n_tags):
        s = rng.uniform(side[0], side[1])
        c = _render_tag(img, rng, s, tilt, 0.0, 255.0, supersample=supersample)
        if c is not None:
            corners.append(c[0])
            payloads.append(c[1])
    return Frame(_finish(img, rng, blur, noise), seed, decimate, corners, payloads)


# ---- the named workloads ---------------------------------------------------

def plain_small(seed=1000):
    """640x480, 8 tags of side 40-120, blur 0.8, noise 4, decimate 2."""
    return make_frame(seed, 640, 480, 8, side=(40, 120), blur=0.8, noise=4.0, decimate=2)


def plain_1080p(seed=2000, decimate=1):
    """1920x1080, 30 tags of side 24-200, 200 shapes + 100 lines, blur 1, noise 6."""
    return make_frame(seed, 1920, 1080, 30, side=(24, 200), n_shapes=200, n_lines=100,
                      blur=1.0, noise=6.0, decimate=decimate)


def stress_12mp(seed=3000):
    """4096x3072, 200 tags of side 16-300, 5000 blobs, a 1024^2 maze, noise 8."""
    return make_frame(seed, 4096, 3072, 200, side=(16, 300), n_blobs=5000, maze=True,
                      blur=0.0, noise=8.0, decimate=1)


BATCH_SIZES = ((640, 480), (1280, 720), (1920, 1080), (2448, 2048), (4096, 3072))


def batch_mixed(seed=4000, n=16):
    """n frames, sizes uniform over BATCH_SIZES, 5-150 tags, 1080p clutter,
    decimate uniform over {1, 2}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    frames = []
    for i in range(n):
        w, h = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        tags = int(rng.integers(5, 151))
        dec = int(rng.integers(1, 3))
        scale = (w * h) / (1920 * 1080)
        frames.append(make_frame(seed + 1 + i, w, h, tags, side=(24, 200),
                                 n_shapes=int(200 * scale), n_lines=int(100 * scale),
                                 blur=1.0, noise=6.0, decimate=dec))
    return frames


def diff_frame(seed=5000, n_tags=24):
    """1280x720 with ground-truth corners, 8x8 supersampled, blur 1.2, noise 3."""
    return make_frame(seed, 1280, 720, n_tags, side=(40, 140), blur=1.2, noise=3.0,
                      supersample=8)


def square_grid(width=4096, height=3072, side=8, pitch=13, margin=3, dark=40, light=200):
    """Dark side x side squares every pitch pixels on a light background:
    one valid cluster per square (73,790 at the defaults, more than the
    reference's u16 cluster ids hold) and a quad for nearly every one."""
    yy = (np.arange(height) - margin) % pitch < side
    xx = (np.arange(width) - margin) % pitch < side
    m = np.outer(yy, xx)
    m[:margin, :] = False
    m[:, :margin] = False
    m[height - pitch:, :] = False
    m[:, width - pitch:] = False
    img = np.full((height, width), light, np.uint8)
    img[m] = dark
    return Frame(img, 0, 1, [])


def decimated_dims(width, height, factor):
    if factor <= 1:
        return width, height
    return 1 + (width - 1) // factor, 1 + (height - 1) // factor
