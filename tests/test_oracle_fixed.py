"""Pins of the oracle's fixed mode (CPU only).

Fixed mode corrects seven behaviours of the reference, so nothing of the
reference can check it; each correction is checked against what the
mathematics fixes instead:

* CCL: scipy.ndimage.label with the 8-neighbourhood on white pixels and the
  4-neighbourhood on black ones (partition, minimum-index roots, sizes).
* boundaries: the records implied by that labelling and the threshold image
  alone (every opposite-colour pixel pair of two large components, column 0
  included, with the reference's bottom-left duplicate rule).
* record decode: the doubled midpoint of the two pixels a record joins, and
  a gradient that points at the white pixel.
* line-fit weight: central differences at the pixel holding the midpoint.
* fit_lines: the reference-faithful work-group emulation at every index but
  global 0, and the cyclic definition at all of them.
* do_indexing: the reference-faithful do_indexing fed each segment's own
  cluster extents.
* end to end: corners within 0.15 px of the ground truth on the
  supersampled frame (compat mode is biased by half a pixel).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import ndimage

from pyatq import records as R
from pyatq import synth
from test_oracle_kats import _cyclic_corners

P = C.c_void_p
WHITE = 0x80000000
# plane -> offset (dx, dy) of the neighbour it tests (src/find_boundaries.cpp:127-166)
PLANES = [(1, 0), (-1, 1), (0, 1), (1, 1)]


@pytest.fixture(scope="module")
def fixed():
    import pyoracle
    return pyoracle.Oracle(fixed=True)


def _components(thr):
    """Per pixel: component id (0 for 127), its minimum pixel index, its size."""
    wl, nw = ndimage.label(thr == 255, structure=np.ones((3, 3), bool))
    bl, nb = ndimage.label(thr == 0, structure=ndimage.generate_binary_structure(2, 1))
    comp = np.where(wl > 0, wl, np.where(bl > 0, bl + nw, 0))
    idx = np.arange(thr.size).reshape(thr.shape)
    mins = np.full(nw + nb + 1, thr.size, np.int64)
    np.minimum.at(mins, comp.ravel(), idx.ravel())
    counts = np.bincount(comp.ravel(), minlength=nw + nb + 1)
    counts[0] = 0
    return comp, mins, counts


def _random_thr(seed, shape, p127):
    rng = np.random.default_rng(seed)
    thr = np.where(rng.random(shape) < 0.5, 0, 255).astype(np.uint8)
    thr[rng.random(shape) < p127] = 127
    return thr


def _frame_thr(oracle, frame):
    img = frame.image
    gray = oracle.decimate(img, frame.decimate) if frame.decimate > 1 else img
    return gray, oracle.threshold(gray)[2]


def _check_labels(thr, labels, sizes):
    comp, mins, counts = _components(thr)
    assert (labels[thr == 127] == 0).all()
    m = thr != 127
    want = (mins[comp[m]] + 1).astype(np.uint32) | np.where(thr[m] == 255, WHITE, 0).astype(np.uint32)
    np.testing.assert_array_equal(labels[m], want)
    assert len(sizes) == thr.size + 1 and sizes[0] == 0
    roots = np.flatnonzero(counts)
    np.testing.assert_array_equal(sizes[mins[roots] + 1], counts[roots])
    assert int(sizes.sum()) == int(m.sum())


@pytest.mark.parametrize("seed,p127", [(0, 0.0), (1, 0.1), (2, 0.4)])
def test_fixed_ccl_is_textbook_labelling_on_random_images(fixed, seed, p127):
    thr = _random_thr(seed, (48, 64), p127)
    labels, sizes = fixed.segmentation(thr)
    _check_labels(thr, labels, sizes)


def test_fixed_ccl_checkerboard_white_diagonal_joins_black_does_not(fixed):
    thr = np.array([[255, 0], [0, 255]], np.uint8)
    labels, sizes = fixed.segmentation(thr)
    assert labels[0, 0] == labels[1, 1] == (WHITE | 1)
    assert labels[0, 1] == 2 and labels[1, 0] == 3
    assert list(sizes) == [0, 2, 1, 1, 0]


@pytest.mark.parametrize("make", [synth.plain_small, synth.diff_frame])
def test_fixed_ccl_on_thresholded_frames(oracle, fixed, make):
    _, thr = _frame_thr(oracle, make())
    labels, sizes = fixed.segmentation(thr)
    _check_labels(thr, labels, sizes)


def _expected_records(thr, labels):
    """Sparse planes implied by the thresholded image and its textbook
    labelling: a record for each pixel pair of opposite colours whose
    components both hold >= 25 pixels; bottom-left is skipped when the left
    and down pixels form such a pair."""
    h, w = thr.shape
    comp, _, counts = _components(thr)
    size = counts[comp]
    big = size >= 25
    n = h * w
    pts = np.zeros(4 * n, np.uint32)
    keys = np.full(4 * n, ~np.uint64(0), np.uint64)
    ys, xs = np.mgrid[0:h, 0:w]

    def pair(dx, dy):
        xn, yn = xs + dx, ys + dy
        inside = (xn >= 0) & (xn < w) & (yn >= 0) & (yn < h)
        xc, yc = np.clip(xn, 0, w - 1), np.clip(yn, 0, h - 1)
        ok = inside & big & big[yc, xc] & (thr.astype(int) + thr[yc, xc] == 255)
        return ok, xc, yc

    # the left pixel and the down pixel: (x-1, y) against (x, y+1)
    down_y = np.clip(ys + 1, 0, h - 1)
    inside = (xs >= 1) & (ys + 1 < h)
    lxx = np.clip(xs - 1, 0, w - 1)
    skip_bl = inside & big[ys, lxx] & big[down_y, xs] & (thr[ys, lxx].astype(int) + thr[down_y, xs] == 255)
    for plane, (dx, dy) in enumerate(PLANES):
        ok, xc, yc = pair(dx, dy)
        if plane == 1:
            ok &= ~skip_bl
        sel = np.flatnonzero(ok.ravel())
        tw = (thr[yc, xc] == 255).ravel()[sel]
        px = (xs.ravel()[sel] | (plane << 14)).astype(np.uint32)
        py = (ys.ravel()[sel] | np.where(tw, 1 << 15, 0)).astype(np.uint32)
        pts[plane * n + sel] = px | (py << 16)
        lv = (labels.ravel()[sel] & 0x7FFFFFFF).astype(np.uint64)
        tv = (labels[yc, xc].ravel()[sel] & 0x7FFFFFFF).astype(np.uint64)
        keys[plane * n + sel] = (np.minimum(lv, tv) << np.uint64(32)) | np.maximum(lv, tv)
    return pts, keys


@pytest.mark.parametrize("case", ["random", "plain_small", "diff_frame"])
def test_fixed_boundaries_are_every_opposite_colour_pair(oracle, fixed, case):
    if case == "random":
        thr = _random_thr(7, (40, 56), 0.05)
        # big components need a coarser pattern than per-pixel noise
        thr = np.kron(thr[:10, :14], np.ones((4, 4), np.uint8))
    else:
        _, thr = _frame_thr(oracle, getattr(synth, case)())
    labels, sizes = fixed.segmentation(thr)
    pts, keys = fixed.find_boundaries(labels, sizes)
    want_p, want_k = _expected_records(thr, labels)
    assert (want_k != ~np.uint64(0)).sum() > 50
    np.testing.assert_array_equal(pts, want_p)
    np.testing.assert_array_equal(keys, want_k)


def test_fixed_boundaries_cover_column_zero(oracle, fixed):
    thr = np.full((40, 40), 255, np.uint8)
    thr[8:32, 0:20] = 0
    n = thr.size
    for o, want in ((fixed, [(2, 7), (3, 7), (2, 31), (3, 31)]), (oracle, [])):
        labels, sizes = o.segmentation(thr)
        _, keys = o.find_boundaries(labels, sizes)
        # column 0 of rows 7 and 31 meets the other colour below and below-right
        got = [(p, i // 40) for p in range(4) for i in range(0, n, 40)
               if keys[p * n + i] != ~np.uint64(0)]
        assert sorted(got) == sorted(want)


def test_fixed_record_decode_is_doubled_midpoint(fixed):
    b = np.zeros(1, R.CLUSTER_BOUNDS)
    for plane, (dx, dy) in enumerate(PLANES):
        for b2w in (0, 1):
            x, y = 37, 21
            px, py = x | (plane << 14), y | (b2w << 15)
            fixed.lib.orc_fx_bounds_from_point(px, py, 5, b.ctypes.data_as(P))
            xn, yn = x + dx, y + dy
            assert (b["x_min"][0], b["y_min"][0]) == (x + xn, y + yn)
            assert (b["x_max"][0], b["y_max"][0]) == (x + xn, y + yn)
            # the gradient points at the white pixel: the neighbour when
            # the record is black-to-white, the local pixel otherwise
            s = 1 if b2w else -1
            assert (b["gx_sum"][0], b["gy_sum"][0]) == (s * dx, s * dy)
            assert b["pxgx_pygy_sum"][0] == (x + xn) * s * dx + (y + yn) * s * dy
            assert (b["start"][0], b["count"][0]) == (5, 1)


def test_fixed_linefit_weight_samples_midpoint_pixel(fixed):
    rng = np.random.default_rng(3)
    h, w = 30, 40
    g = rng.integers(0, 256, (h, w), dtype=np.uint8)
    out = np.zeros(1, R.LINE_FIT_POINT)
    checked = 0
    for plane, (dx, dy) in enumerate(PLANES):
        for x, y in [(5, 17), (20, 3), (33, 25), (1, 1), (0, 10), (38, 28)]:
            if not (0 <= x + dx < w and 0 <= y + dy < h):
                continue
            px, py = x | (plane << 14), y
            fixed.lib.orc_fx_initial_linefit(px, py, w, h, g.ctypes.data_as(P), out.ctypes.data_as(P))
            rx, ry = 2 * x + dx + 1, 2 * y + dy + 1
            cx, cy = rx // 2, ry // 2
            W = 1
            if 0 < cx < w - 1 and 0 < cy < h - 1:
                gx = int(g[cy, cx + 1]) - int(g[cy, cx - 1])
                gy = int(g[cy + 1, cx]) - int(g[cy - 1, cx])
                W += math.isqrt(gx * gx + gy * gy)
                checked += cx != cy
            r = out[0]
            assert (r["W"], r["Mx"], r["My"]) == (W, W * rx, W * ry)
            assert (r["Mxx"], r["Myy"], r["Mxy"]) == (W * rx * rx, W * ry * ry, W * rx * ry)
    assert checked >= 8


@pytest.mark.parametrize("seed", [0, 1])
def test_fixed_fit_lines_is_cyclic_definition_everywhere(fixed, seed):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(24, 70, 5)
    idx = np.repeat(np.arange(len(sizes)), sizes).astype(np.uint32)
    ext = np.zeros(len(sizes), R.CLUSTER_EXTENTS)
    ext["count"] = sizes
    ext["start"] = np.r_[0, np.cumsum(sizes)[:-1]]
    raw = np.zeros(len(idx), R.LINE_FIT_POINT)
    for s, n in zip(ext["start"], sizes):
        th = np.linspace(0, 2 * np.pi, n, endpoint=False) + 0.3
        r = 20 + 8 * np.abs(np.cos(2 * th))
        xs = np.rint(100 + r * np.cos(th) + rng.normal(0, 0.7, n)).astype(np.int64)
        ys = np.rint(100 + r * np.sin(th) + rng.normal(0, 0.7, n)).astype(np.int64)
        ws = rng.integers(1, 40, n)
        rx, ry = 2 * xs + 1, 2 * ys + 1
        sl = slice(int(s), int(s) + int(n))
        raw["W"][sl] = ws
        raw["Mx"][sl], raw["My"][sl] = ws * rx, ws * ry
        raw["Mxx"][sl], raw["Myy"][sl], raw["Mxy"][sl] = ws * rx * rx, ws * ry * ry, ws * rx * ry
    lfp = fixed.linefit_prefix(idx, raw)
    got = fixed.fit_lines(lfp, idx, ext)
    want = _cyclic_corners(fixed, lfp, idx, ext, first=0)
    assert got.tobytes() == want.tobytes()


@pytest.fixture(scope="module")
def fixed_1080p(fixed):
    fr = synth.plain_1080p()
    return fixed.pipeline(fr.image, fr.decimate)


def test_fixed_fit_lines_equals_emulation_except_index_zero(oracle, fixed, fixed_1080p):
    o = fixed_1080p
    compat = oracle.fit_lines(o["line_fit"], o["cluster_index"], o["extents"])
    got = o["sparse_corners"]
    assert got[1:].tobytes() == compat[1:].tobytes()
    assert compat[0]["error"] == 0


def test_fixed_do_indexing_fits_each_segment_to_its_own_cluster(oracle, fixed, fixed_1080p):
    o = fixed_1080p
    seg, cc = o["peak_extents"], o["corners"]
    own = cc["cluster_index"][seg["start"]]
    want, _ = oracle.do_indexing(seg, cc, o["line_fit"], o["extents"][own])
    assert o["fitted"].tobytes() == want.tobytes()
    # with clusters 0, 5 and 17 stripped of their corners, segment e no
    # longer belongs to cluster e
    sparse = o["sparse_corners"].copy()
    sparse["error"][np.isin(sparse["cluster_index"], [0, 5, 17])] = 0
    cc2, seg2 = fixed.corner_segments(sparse)
    own2 = cc2["cluster_index"][seg2["start"]]
    assert (own2 != np.arange(len(seg2))).sum() > len(seg2) // 2
    got, _ = fixed.do_indexing(seg2, cc2, o["line_fit"], o["extents"])
    want, _ = oracle.do_indexing(seg2, cc2, o["line_fit"], o["extents"][own2])
    assert (got["indices"] != 0).any()
    assert got.tobytes() == want.tobytes()


def _mean_corner_error(quads, truth, scale):
    errs = []
    for t in truth:
        t = np.asarray(t, np.float64)
        best = np.inf
        for c in quads["corners"]:
            c = np.asarray(c, np.float64) * scale
            for r in range(4):
                for cc in (c, c[::-1]):
                    best = min(best, float(np.sqrt(((np.roll(cc, r, 0) - t) ** 2).sum(1)).mean()))
        errs.append(best)
    errs = np.array(errs)
    return errs[errs < 2.0]


def test_fixed_corners_are_unbiased_on_supersampled_frame(oracle, fixed):
    fr = synth.diff_frame()
    qc, _ = oracle.detect(fr.image, 1)
    qf, _ = fixed.detect(fr.image, 1)
    ec = _mean_corner_error(qc, fr.corners, 1)
    ef = _mean_corner_error(qf, fr.corners, 1)
    assert len(ef) >= len(ec) >= 8
    assert ef.mean() < 0.15 < 0.4 < ec.mean()


def test_fixed_pipeline_equals_fixed_detect(fixed, fixed_1080p):
    fr = synth.plain_1080p()
    q, st = fixed.detect(fr.image, fr.decimate)
    assert q.tobytes() == fixed_1080p["quads"].tobytes()
    assert st["n_clusters"] == len(fixed_1080p["bounds"])
