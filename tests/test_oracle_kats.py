"""Known answers, closed forms and independent restatements that pin the
oracle's stages (CPU only).

* decimate: numpy slicing.
* threshold: hand-derived tiles (flat, step, the 5-level contrast rule, the
  integer midpoint, edge-clipped dilation).
* segmentation: hand-derived labels including the colour-agnostic neighbour
  mask that bridges two white pixels which are not 8-connected.
* find_boundaries: the 8x32 work-group emulation against a per-pixel
  statement of the rule, plus hand-counted records of a square.
* fit_line: the eigen-decomposition of the weighted covariance in float64.
* get_moment: direct sums over cyclic ranges.
* fit_lines: the work-group emulation against the per-cluster cyclic
  definition evaluated here in float32.
"""
import ctypes as C

import numpy as np
import pytest

from pyatq import records as R

P = C.c_void_p


def test_decimate_is_point_subsample(oracle):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (75, 103), dtype=np.uint8)
    for f in (2, 3, 4):
        np.testing.assert_array_equal(oracle.decimate(img, f), img[::f, ::f])


def test_threshold_flat_is_127(oracle):
    _, _, thr = oracle.threshold(np.full((16, 24), 77, np.uint8))
    assert (thr == 127).all()




def test_threshold_step_and_contrast_rule(oracle):
    img = np.zeros((8, 8), np.uint8)
    img[:, 4:] = 100
    ext, fil, thr = oracle.threshold(img)
    assert (ext["min"] == [[0, 100], [0, 100]]).all() and (ext["max"] == [[0, 100], [0, 100]]).all()
    assert (fil["min"] == 0).all() and (fil["max"] == 100).all()
    want = np.zeros((8, 8), np.uint8)
    want[:, 4:] = 255  # 100 > 0 + 100 // 2
    np.testing.assert_array_equal(thr, want)
    # max - min == 4 stays grey, == 5 thresholds at min + 5 // 2 = 12
    for hi, expect in ((14, None), (15, 255)):
        img = np.full((8, 8), 10, np.uint8)
        img[0, 0] = hi
        _, _, thr = oracle.threshold(img)
        if expect is None:
            assert (thr == 127).all()
        else:
            assert thr[0, 0] == 255 and (thr.reshape(-1)[1:] == 0).all()
    # integer midpoint: min 3, max 10 -> comp 6: 6 -> 0, 7 -> 255
    img = np.full((4, 4), 3, np.uint8)
    img[0, 0], img[0, 1], img[0, 2] = 10, 6, 7
    _, _, thr = oracle.threshold(img)
    assert (thr[0, 0], thr[0, 1], thr[0, 2], thr[1, 1]) == (255, 0, 255, 0)


def test_threshold_dilation_is_clipped_3x3(oracle):
    img = np.zeros((4, 16), np.uint8)
    img[:, 12:] = 200          # only tile 3 (of 4 in a row) is bright
    ext, fil, thr = oracle.threshold(img)
    assert list(fil["max"][0]) == [0, 0, 200, 200]
    assert (thr[:, :8] == 127).all()
    assert (thr[:, 8:12] == 0).all() and (thr[:, 12:] == 255).all()


def _seg(oracle, thr):
    labels, sizes = oracle.segmentation(np.ascontiguousarray(thr, np.uint8))
    return labels, sizes


def test_segmentation_uniform_images(oracle):
    lab, siz = _seg(oracle, np.full((8, 8), 255))
    assert (lab == 0x80000000).all() and siz[0] == 64
    lab, siz = _seg(oracle, np.full((8, 8), 0))
    assert (lab == 1).all() and siz[1] == 64          # black root: block 0's second node
    lab, siz = _seg(oracle, np.full((8, 8), 127))
    assert (lab == 0).all() and siz[0] == 0


def test_segmentation_labels_are_min_node(oracle):
    t = np.full((8, 8), 127)
    t[0:2, 0:2] = 255
    t[6:8, 6:8] = 255
    lab, siz = _seg(oracle, t)
    assert (lab[0:2, 0:2] == 0x80000000).all()
    assert (lab[6:8, 6:8] == (0x80000000 | (6 * 8 + 6))).all()
    assert siz[0] == 4 and siz[54] == 4
    # diagonal 8-connectivity across blocks
    t = np.full((8, 8), 127)
    t[1, 1] = 255
    t[2, 2] = 255
    lab, siz = _seg(oracle, t)
    assert lab[1, 1] == lab[2, 2] == 0x80000000 and siz[0] == 2


def test_segmentation_colour_agnostic_bridge(oracle):
    """Block (1,1) holds only black pixels, yet its neighbour mask (set by any
    non-127 pixel) tests P = (1,1) and R = (4,1); both are white, so its empty
    white set links block (0,0)'s and block (2,0)'s white sets: two white
    pixels three columns apart share one label (textbook CCL separates them)."""
    t = np.full((8, 8), 127)
    t[1, 1] = 255
    t[1, 4] = 255
    t[2, 2] = 0
    t[2, 3] = 0
    lab, siz = _seg(oracle, t)
    assert lab[1, 1] == 0x80000000 and lab[1, 4] == 0x80000000
    assert siz[0] == 2
    assert lab[2, 2] == lab[2, 3] == 2 * 8 + 2 + 1 and siz[19] == 2


def _boundaries_per_pixel(labels, sizes, min_size=25):
    """The boundary rule stated per pixel (no work-group tiling)."""
    h, w = labels.shape
    pts = np.zeros(4 * w * h, np.uint32)
    keys = np.full(4 * w * h, np.uint64(0xFFFFFFFFFFFFFFFF))

    def lab(x, y):
        return int(labels[y, x]) if (0 <= x < w and 0 <= y < h) else 0

    for y in range(h):
        for x in range(1, w):
            L = lab(x, y)
            lv, wl = L & 0x7FFFFFFF, L & 0x80000000
            left, down = lab(x - 1, y), lab(x, y + 1)
            dup = ((left & 0x80000000) != (down & 0x80000000) and sizes[left & 0x7FFFFFFF] >= min_size
                   and sizes[down & 0x7FFFFFFF] >= min_size)
            for plane, (dx, dy) in enumerate(((1, 0), (-1, 1), (0, 1), (1, 1))):
                if plane == 1 and dup:
                    continue
                T = lab(x + dx, y + dy)
                wt, tv = T & 0x80000000, T & 0x7FFFFFFF
                if wt == wl or sizes[lv] < min_size or sizes[tv] < min_size:
                    continue
                px = x | (plane << 14)
                py = (0x8000 | y) if wt > wl else y
                i = plane * w * h + y * w + x
                pts[i] = px | (py << 16)
                keys[i] = (min(lv, tv) << 32) | max(lv, tv)
    return pts, keys


@pytest.mark.parametrize("seed,shape", [(0, (24, 40)), (1, (36, 64)), (2, (16, 96))])
def test_find_boundaries_emulation_matches_per_pixel_rule(oracle, seed, shape):
    rng = np.random.default_rng(seed)
    h, w = shape
    img = np.kron(rng.integers(0, 3, (h // 4, w // 4)), np.ones((4, 4))).astype(np.uint8)
    thr = np.choose(img, [0, 127, 255]).astype(np.uint8)
    thr[rng.random((h, w)) < 0.05] = 127
    labels, sizes = oracle.segmentation(thr)
    pts, keys = oracle.find_boundaries(labels, sizes)
    want_p, want_k = _boundaries_per_pixel(labels, sizes)
    assert pts.any()
    np.testing.assert_array_equal(pts, want_p)
    np.testing.assert_array_equal(keys, want_k)


def test_find_boundaries_square_counts(oracle):
    t = np.zeros((16, 16), np.uint8)
    t[4:12, 4:12] = 255
    labels, sizes = oracle.segmentation(t)
    assert sizes[68] == 64 and sizes[1] == 192
    pts, keys = oracle.find_boundaries(labels, sizes)
    n = 16 * 16
    planes = (pts.reshape(4, n) != 0).sum(1)
    assert planes[0] == 16 and planes[2] == 16          # RIGHT and BOTTOM edges
    k = keys[pts != 0]
    assert (k == ((1 << 32) | 68)).all()
    # column 0 never emits (it is a halo column in every work-group)
    t2 = np.zeros((16, 64), np.uint8)
    t2[:, :8] = 255
    l2, s2 = oracle.segmentation(t2)
    p2, _ = oracle.find_boundaries(l2, s2)
    xs = p2[p2 != 0] & 0x3FFF
    assert xs.min() >= 1


def _linefit_moment(xs, ys, ws):
    m = np.zeros(1, R.LINE_FIT_POINT)
    rx, ry = 2 * xs.astype(np.int64), 2 * ys.astype(np.int64)
    m["Mx"], m["My"] = (ws * rx).sum(), (ws * ry).sum()
    m["Mxx"], m["Myy"], m["Mxy"] = (ws * rx * rx).sum(), (ws * ry * ry).sum(), (ws * rx * ry).sum()
    m["W"] = ws.sum()
    return m


def test_fit_line_is_covariance_eigen(oracle):
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = int(rng.integers(5, 60))
        ang = rng.uniform(0, np.pi)
        t = rng.uniform(-30, 30, n)
        xs = np.rint(500 + t * np.cos(ang) + rng.normal(0, 1.5, n)).astype(np.int64)
        ys = np.rint(400 + t * np.sin(ang) + rng.normal(0, 1.5, n)).astype(np.int64)
        ws = rng.integers(1, 50, n).astype(np.int64)
        m = _linefit_moment(xs, ys, ws)
        line = np.zeros(4, np.float32)
        err, mse = C.c_float(), C.c_float()
        oracle.lib.orc_fit_line(m.ctypes.data_as(P), n, line.ctypes.data_as(P), C.byref(err),
                                C.byref(mse))
        W = ws.sum()
        mx, my = (ws * xs).sum() / W, (ws * ys).sum() / W
        cov = np.array([[(ws * (xs - mx) ** 2).sum(), (ws * (xs - mx) * (ys - my)).sum()],
                        [(ws * (xs - mx) * (ys - my)).sum(), (ws * (ys - my) ** 2).sum()]]) / W
        lam, vec = np.linalg.eigh(cov)
        assert abs(mse.value - lam[0]) <= 1e-3 * max(1.0, lam[1])
        assert abs(err.value - n * mse.value) <= 1e-6 * max(1.0, abs(err.value))
        assert abs(line[0] - mx) < 1e-3 and abs(line[1] - my) < 1e-3
        # (line[2], line[3]) is the unit normal: the minor-eigenvalue eigenvector
        assert abs(abs(line[2] * vec[0, 0] + line[3] * vec[1, 0]) - 1) < 2e-3


def test_fit_line_horizontal_closed_form(oracle):
    xs = np.arange(10, 30, dtype=np.int64)
    ys = np.full(20, 7, np.int64)
    m = _linefit_moment(xs, ys, np.ones(20, np.int64))
    line = np.zeros(4, np.float32)
    err, mse = C.c_float(), C.c_float()
    oracle.lib.orc_fit_line(m.ctypes.data_as(P), 20, line.ctypes.data_as(P), C.byref(err), C.byref(mse))
    assert mse.value == 0.0 and err.value == 0.0
    np.testing.assert_array_equal(line, np.array([19.5, 7.0, 0.0, -1.0], np.float32))


def _prefix(raw):
    out = raw.copy()
    for f in ("Mx", "My", "Mxx", "Myy", "Mxy", "W"):
        out[f] = np.cumsum(raw[f])
    return out


def test_get_moment_cyclic_sums(oracle):
    rng = np.random.default_rng(3)
    n = 17
    raw = np.zeros(n, R.LINE_FIT_POINT)
    for f in ("Mx", "My", "Mxx", "Myy", "Mxy", "W"):
        raw[f] = rng.integers(0, 1000, n)
    pre = _prefix(raw)
    for i0 in range(n):
        for i1 in range(n):
            out = np.zeros(1, R.LINE_FIT_POINT)
            tot = C.c_int64()
            oracle.lib.orc_get_moment(pre.ctypes.data_as(P), n, i0, i1, out.ctypes.data_as(P),
                                      C.byref(tot))
            if i0 == i1:
                assert tot.value == 0 and out["W"][0] == 0
                continue
            idx = np.arange(i0, i1 + 1) if i0 < i1 else np.r_[np.arange(i0, n), np.arange(0, i1 + 1)]
            assert tot.value == len(idx)
            for f in ("Mx", "My", "Mxx", "Myy", "Mxy", "W"):
                assert out[f][0] == raw[f][idx].sum()


# src/fit_quad.cpp:11-13, as hex float literals
FILTER = np.array([float.fromhex(h) for h in ("0x1.6c0504p-7", "0x1.152aaap-3", "0x1.368b3p-1",
                                              "0x1p+0", "0x1.368b3p-1", "0x1.152aaap-3",
                                              "0x1.6c0504p-7")], np.float32)


def test_filter_taps_are_float_gaussian():
    d = np.arange(-3, 4)
    np.testing.assert_array_equal(FILTER, np.exp(-(d * d) / 2.0).astype(np.float32))


def _cyclic_corners(oracle, lfp, idx, ext, first=1):
    """The per-cluster cyclic window / filter / strict maximum, in float32,
    for the points at global index >= first."""
    n_all = len(lfp)
    out = np.zeros(n_all, R.CORNER)
    for c, (s, n) in enumerate(zip(ext["start"], ext["count"])):
        s, n = int(s), int(n)
        k = min(20, n // 12)
        seg = np.ascontiguousarray(lfp[s:s + n])
        e = np.zeros(n, np.float32)
        for p in range(n):
            mom = np.zeros(1, R.LINE_FIT_POINT)
            tot = C.c_int64()
            oracle.lib.orc_get_moment(seg.ctypes.data_as(P), n, (p - k) % n, (p + k) % n,
                                      mom.ctypes.data_as(P), C.byref(tot))
            err, mse = C.c_float(), C.c_float()
            oracle.lib.orc_fit_line(mom.ctypes.data_as(P), tot.value, None, C.byref(err), C.byref(mse))
            e[p] = err.value
        f = np.zeros(n, np.float32)
        for p in range(n):
            acc = np.float32(0)
            for t in range(7):
                acc = np.float32(acc + np.float32(e[(p + t - 3) % n] * FILTER[t]))
            f[p] = acc
        for p in range(n):
            x = s + p
            if x >= first and f[p] > f[(p - 1) % n] and f[p] > f[(p + 1) % n]:
                out[x] = (p, c, f[p])
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fit_lines_emulation_equals_cyclic_definition(oracle, seed):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(24, 70, 6)
    idx = np.repeat(np.arange(len(sizes)), sizes).astype(np.uint32)
    ext = np.zeros(len(sizes), R.CLUSTER_EXTENTS)
    ext["count"] = sizes
    ext["start"] = np.r_[0, np.cumsum(sizes)[:-1]]
    raw = np.zeros(len(idx), R.LINE_FIT_POINT)
    for c, (s, n) in enumerate(zip(ext["start"], sizes)):
        th = np.linspace(0, 2 * np.pi, n, endpoint=False)
        r = 20 + 8 * np.abs(np.cos(2 * th))  # a rounded square-ish contour
        xs = np.rint(100 + r * np.cos(th) + rng.normal(0, 0.7, n)).astype(np.int64)
        ys = np.rint(100 + r * np.sin(th) + rng.normal(0, 0.7, n)).astype(np.int64)
        ws = rng.integers(1, 40, n)
        rx, ry = 2 * xs + 2, 2 * ys + 1
        sl = slice(int(s), int(s) + int(n))
        raw["W"][sl] = ws
        raw["Mx"][sl], raw["My"][sl] = ws * rx, ws * ry
        raw["Mxx"][sl], raw["Myy"][sl], raw["Mxy"][sl] = ws * rx * rx, ws * ry * ry, ws * rx * ry
    lfp = oracle.linefit_prefix(idx, raw)
    got = oracle.fit_lines(lfp, idx, ext)
    want = _cyclic_corners(oracle, lfp, idx, ext)
    assert (want["error"] != 0).sum() > 4
    assert got.tobytes() == want.tobytes()
