"""CPU restatement of the tag-payload decode (TEST INFRASTRUCTURE ONLY).

The reference decodes every quad with the apriltag library's
quad_decode_task (src/quad_decoding.cpp:46-75, called per quad from
decode_quads, src/quad_decoding.cpp:77-128, after process_fitted_quads,
src/main.cpp:1164). That library is a dependency fetched at build time
(carbotaniuman/apriltag @ a8b2f9639765162f2ac12a608d785ceb97deee6b,
CMakeLists.txt:11-17) and is absent from the reference checkout, as is its
tag36h11 family (code table and bit layout). This file restates the
table-free part of apriltag 3's published quad_decode (apriltag.c):

* the homography from the tag square [-1, 1]^2 to the quad, corner i at
  (-1,-1), (1,-1), (1,1), (-1,1) (quad_update_homographies), solved here
  as the exact 8x8 linear system of the four correspondences;
* the white and black border gray models: 8 lines of width_at_border
  samples just outside and just inside the black border, each pixel read
  at the truncated projected position, each model the least-squares plane
  I = c0 x + c1 y + c2 over its samples (graymodel_add / graymodel_solve);
* the polarity test (white - black at the centre must not be negative
  unless the family has a reversed border);
* one bilinear sample per bit cell (value_for_pixel: pixel centres at
  +0.5, -1 outside the image), minus the mean of the two models there,
  stored as float in a total_width^2 grid, then sharpened by
  v += decode_sharpening * (4 v - the 4 neighbours) (sharpen, zero outside
  the grid);
* the code word, bit i (the family's bit order, most significant first)
  set where the sharpened value is > 0, and the decision margin
  min(mean white margin, mean black margin) with counts starting at 1.

The code-table lookup (quick_decode_codeword) is replaced by `match`: the
nearest code of a caller-supplied table over the four 90-degree rotations.
Parity with the fork's quad_decode_task itself is unpinned (the fork is not
here to run); the pins are closed forms (test_decode.py).
"""
import numpy as np


def tag_corners():
    """Tag-square coordinates of quad corners 0..3."""
    return np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])


def homography(corners):
    """3x3 H (H[2,2] = 1) with H (tx, ty, 1) ~ (px, py, 1) for the four
    corner correspondences (quad_update_homographies)."""
    src = tag_corners()
    dst = np.asarray(corners, np.float64).reshape(4, 2)
    a = np.zeros((8, 8))
    b = np.zeros(8)
    for i in range(4):
        x, y = src[i]
        u, v = dst[i]
        a[2 * i] = [x, y, 1, 0, 0, 0, -x * u, -y * u]
        a[2 * i + 1] = [0, 0, 0, x, y, 1, -x * v, -y * v]
        b[2 * i], b[2 * i + 1] = u, v
    h = np.linalg.solve(a, b)
    return np.append(h, 1.0).reshape(3, 3)


def project(h, x, y):
    d = h[2, 0] * x + h[2, 1] * y + h[2, 2]
    return (h[0, 0] * x + h[0, 1] * y + h[0, 2]) / d, (h[1, 0] * x + h[1, 1] * y + h[1, 2]) / d


def _plane(samples):
    """Least-squares plane c0 x + c1 y + c2 through (x, y, v) samples, or
    None when they do not span a plane (fewer than 3 samples off one line;
    the normal matrix's smallest eigenvalue at or below 1e-9)."""
    s = np.asarray(samples, np.float64).reshape(-1, 3)
    if len(s) < 3:
        return None
    a = np.c_[s[:, 0], s[:, 1], np.ones(len(s))]
    n = a.T @ a
    if np.linalg.eigvalsh(n)[0] <= 1e-9:
        return None
    return np.linalg.solve(n, a.T @ s[:, 2])


def value_for_pixel(gray, px, py):
    h, w = gray.shape
    if not (np.isfinite(px) and np.isfinite(py)):
        return -1.0
    x1, x2 = int(np.floor(px - 0.5)), int(np.ceil(px - 0.5))
    y1, y2 = int(np.floor(py - 0.5)), int(np.ceil(py - 0.5))
    if x1 < 0 or x2 >= w or y1 < 0 or y2 >= h:
        return -1.0
    fx, fy = px - 0.5 - x1, py - 0.5 - y1
    g = gray.astype(np.float64)
    return (g[y1, x1] * (1 - fx) * (1 - fy) + g[y1, x2] * fx * (1 - fy) +
            g[y2, x1] * (1 - fx) * fy + g[y2, x2] * fx * fy)


def border_models(gray, h, wb):
    """(white, black) plane coefficients from the 8 border sample lines."""
    pat = [(-0.5, 0.5, 0, 1, 1), (0.5, 0.5, 0, 1, 0), (wb + 0.5, 0.5, 0, 1, 1),
           (wb - 0.5, 0.5, 0, 1, 0), (0.5, -0.5, 1, 0, 1), (0.5, 0.5, 1, 0, 0),
           (0.5, wb + 0.5, 1, 0, 1), (0.5, wb - 0.5, 1, 0, 0)]
    hh, ww = gray.shape
    white, black = [], []
    for x0, y0, dx, dy, is_white in pat:
        for i in range(wb):
            tx = 2 * ((x0 + i * dx) / wb - 0.5)
            ty = 2 * ((y0 + i * dy) / wb - 0.5)
            px, py = project(h, tx, ty)
            if not (np.isfinite(px) and np.isfinite(py) and abs(px) < 2e9 and abs(py) < 2e9):
                continue
            ix, iy = int(px), int(py)  # C truncation toward zero
            if ix < 0 or iy < 0 or ix >= ww or iy >= hh:
                continue
            (white if is_white else black).append((tx, ty, float(gray[iy, ix])))
    return white, black


def decode(gray, corners, layout):
    """(code, decision margin, H) of one quad; margin -1 and code 0 when the
    border polarity test fails."""
    wb, tw = layout["width_at_border"], layout["total_width"]
    h = homography(corners)
    white, black = border_models(gray, h, wb)
    cw, cb = _plane(white), _plane(black)
    if cw is None or cb is None:  # too few border samples in the image
        return 0, -1.0, h
    if ((cw[2] - cb[2]) < 0) != bool(layout.get("reversed_border", 0)):
        return 0, -1.0, h
    vals = np.zeros((tw, tw), np.float32)
    mn = (wb - tw) // 2
    for bx, by in zip(layout["bit_x"], layout["bit_y"]):
        tx = 2 * ((bx + 0.5) / wb - 0.5)
        ty = 2 * ((by + 0.5) / wb - 0.5)
        px, py = project(h, tx, ty)
        v = value_for_pixel(gray, px, py)
        if v == -1:
            continue
        th = (cb @ [tx, ty, 1.0] + cw @ [tx, ty, 1.0]) / 2.0
        vals[by - mn, bx - mn] = np.float32(v - th)
    sharp = np.zeros_like(vals)
    k = np.array([[0, -1, 0], [-1, 4, -1], [0, -1, 0]], np.float32)
    for y in range(tw):
        for x in range(tw):
            acc = np.float32(0)
            for i in range(3):
                for j in range(3):
                    yy, xx = y + i - 1, x + j - 1
                    if 0 <= yy < tw and 0 <= xx < tw:
                        acc = np.float32(acc + vals[yy, xx] * k[i, j])
            sharp[y, x] = acc
    sh = float(layout.get("decode_sharpening", 0.25))
    vals = (vals.astype(np.float64) + sh * sharp.astype(np.float64)).astype(np.float32)
    code = 0
    ws, bs, wc, bc = np.float32(0), np.float32(0), np.float32(1), np.float32(1)
    for bx, by in zip(layout["bit_x"], layout["bit_y"]):
        v = vals[by - mn, bx - mn]
        code <<= 1
        if v > 0:
            ws, wc = np.float32(ws + v), np.float32(wc + 1)
            code |= 1
        else:
            bs, bc = np.float32(bs - v), np.float32(bc + 1)
    return code, float(min(ws / wc, bs / bc)), h


def rotation_permutations(layout):
    """perm[r][i]: the bit of the code read r quarter turns later (the cell
    (x, y) of bit i, border coordinates 0 .. wb-1, taken to (wb-1-y, x)
    r times) that bit i's value lands on; None when the layout's cells are
    not closed under the quarter turn."""
    wb = layout["width_at_border"]
    cells = {(x, y): i for i, (x, y) in enumerate(zip(layout["bit_x"], layout["bit_y"]))}
    out = []
    for r in range(4):
        perm = []
        for x, y in zip(layout["bit_x"], layout["bit_y"]):
            for _ in range(r):
                x, y = wb - 1 - y, x
            if (x, y) not in cells:
                return None
            perm.append(cells[(x, y)])
        out.append(perm)
    return out


def rotate(code, layout, r):
    n = len(layout["bit_x"])
    perm = rotation_permutations(layout)[r]
    out = 0
    for i in range(n):
        if (code >> (n - 1 - i)) & 1:
            out |= 1 << (n - 1 - perm[i])
    return out


def match(code, table, layout, max_hamming):
    """(id, hamming, rotation) of the nearest table code over the four
    rotations of `code` (ties: lower rotation, then lower id); id -1 when
    none is within max_hamming."""
    best = (-1, max_hamming + 1, 0)
    for r in range(4):
        c = rotate(code, layout, r)
        for i, t in enumerate(table):
            d = bin(int(c) ^ int(t)).count("1")
            if d < best[1]:
                best = (i, d, r)
    return best if best[0] >= 0 else (-1, 0, 0)


def row_major_layout(n=6, border=1, wb=8, tw=10):
    """The 6x6 payload of the synthetic tags, bits row by row (bit 0 = the
    top-left cell, most significant), cells border .. border+n-1."""
    xs, ys = [], []
    for y in range(n):
        for x in range(n):
            xs.append(border + x)
            ys.append(border + y)
    return {"width_at_border": wb, "total_width": tw, "reversed_border": 0,
            "bit_x": xs, "bit_y": ys, "decode_sharpening": 0.25}


def payload_code(cells6):
    """Code word of a 6x6 0/1 payload (row-major, white = 1)."""
    c = 0
    for v in np.asarray(cells6).reshape(-1):
        c = (c << 1) | int(v > 0.5)
    return c
