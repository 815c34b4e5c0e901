"""Independent statements that pin the oracle functions the reference's own
code cannot (CPU only).

* orc_segmentation: a breadth-first labelling (scipy connected components)
  of an explicit edge list written from src/segmentation.cpp:139-255 (the
  colour-agnostic neighbour mask, its image-edge masks, the six probes and
  the neighbour block each probe names), on 12,000 random ternary images of
  sizes down to 4 x 4 (every edge-mask case), labels = minimum node id.
* orc_do_indexing's selection rule (src/fit_quad.cpp:445-519): brute force
  over every 4-subset of a segment's first min(count, 10) corners,
  enumerated in the textbook combinatorial number system, with the
  per-segment fits taken from the oracle's own get_moment/fit_line (pinned
  in test_oracle_kats.py): reject any fit with mse > 10 and lines 0-1 only
  at |cos| > cos 10 deg, then take the SMALLEST combination number whose
  error is within 0.1 of the minimum (not the argmin).
* orc_process_quads (src/main.cpp:1035-1162): closed forms. Four exactly
  collinear point sets on the sides of a known quad give its corners;
  reversed winding, parallel consecutive sides, areas under 0.95 * 8^2 and
  angles under 10 degrees are rejected; never-written fits are skipped.
* the float64 twin of the quad corners (pyoracle.quad_corners_fp64): the
  float32 corners are within 1e-3 px of it on the plain frames.
"""
import ctypes as C
import math
from itertools import combinations

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from pyatq import records as R

P = C.c_void_p


# ---------------------------------------------------------------- segmentation
# (pixel dy, pixel dx, mask bit, neighbour block dby, dbx): the six
# handle_pixel_test calls of src/segmentation.cpp:242-247 with the block each
# label offset names (:197-236): P -2w-2, Q -2w, R -2w+2, S -2.
EDGE_PROBES = [(-1, -1, 0, -1, -1), (-1, 0, 1, -1, 0), (-1, 1, 2, -1, 0),
               (-1, 2, 3, -1, 1), (0, -1, 4, 0, -1), (1, -1, 8, 0, -1)]


def bke_reference_labels(thr):
    """Labels and sizes of the reference's block graph by connected
    components over its explicit edge list."""
    H, W = thr.shape
    bh, bw = H // 2, W // 2
    t = thr.astype(np.int64)
    tl, tr, bl, br = t[0::2, 0::2], t[0::2, 1::2], t[1::2, 0::2], t[1::2, 1::2]
    y = 2 * np.arange(bh)[:, None] + np.zeros((1, bw), np.int64)
    x = 2 * np.arange(bw)[None, :] + np.zeros((bh, 1), np.int64)
    coloured = lambda v: (v == 0) | (v == 255)  # noqa: E731
    m = (np.where(coloured(tl), 0x777, 0) | np.where(coloured(tr), 0x777 << 1, 0)
         | np.where(coloured(bl), 0x777 << 4, 0))
    m = np.where(x == 0, m & 0xEEEE, np.where(x + 2 >= W, m & 0x7777, m))
    m = np.where(y == 0, m & 0xFFF0, np.where(y + 2 >= H, m & 0x0FFF, m))
    pad = np.full((H + 2, W + 3), 127, np.int64)
    pad[1:H + 1, 1:W + 1] = t
    node = lambda by, bx, c: 2 * (by * bw + bx) + c  # noqa: E731
    bys, bxs = np.indices((bh, bw))
    src, dst = [], []
    for dy, dx, bit, dby, dbx in EDGE_PROBES:
        pv = pad[1 + y + dy, 1 + x + dx]
        tested = ((m >> bit) & 1).astype(bool)
        for c, colour in ((0, 255), (1, 0)):
            sel = tested & (pv == colour)
            src.append(node(bys[sel], bxs[sel], c))
            dst.append(node(bys[sel] + dby, bxs[sel] + dbx, c))
    src = np.concatenate(src) if src else np.zeros(0, np.int64)
    dst = np.concatenate(dst) if dst else np.zeros(0, np.int64)
    n = 2 * bh * bw
    g = coo_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
    ncomp, comp = connected_components(g, directed=False)
    ids = np.arange(n)
    gid = 2 * (ids // 2 // bw) * W + 2 * (ids // 2 % bw) + ids % 2  # TL pixel index + colour
    root = np.full(ncomp, np.iinfo(np.int64).max)
    np.minimum.at(root, comp, gid)
    node_root = root[comp].reshape(bh, bw, 2)
    labels = np.zeros((H, W), np.uint32)
    sizes = {}
    for oy in (0, 1):
        for ox in (0, 1):
            v = t[oy::2, ox::2]
            lw = (node_root[..., 0] | 0x80000000).astype(np.uint32)
            lb = node_root[..., 1].astype(np.uint32)
            labels[oy::2, ox::2] = np.where(v == 255, lw, np.where(v == 0, lb, 0))
            for c, colour in ((0, 255), (1, 0)):
                r, k = np.unique(node_root[..., c][v == colour], return_counts=True)
                for ri, ki in zip(r.tolist(), k.tolist()):
                    sizes[ri] = sizes.get(ri, 0) + ki
    return labels, sizes


def test_segmentation_equals_edge_set_components(oracle):
    rng = np.random.default_rng(2024)
    dims = [(4, 4), (4, 8), (8, 4), (8, 8), (6, 10), (10, 6), (12, 8), (8, 16), (16, 16),
            (14, 10), (20, 12), (24, 24)]
    n_images = 0
    for i in range(12000):
        h, w = dims[i % len(dims)]
        p127 = (0.0, 0.1, 0.3, 0.6, 0.9)[(i // len(dims)) % 5]
        pw = rng.uniform(0.1, 0.9)
        u = rng.random((h, w))
        thr = np.where(u < p127, 127, np.where(rng.random((h, w)) < pw, 255, 0)).astype(np.uint8)
        lab, siz = oracle.segmentation(thr)
        want, want_sizes = bke_reference_labels(thr)
        np.testing.assert_array_equal(lab, want, err_msg=f"image {i} {h}x{w}")
        for r, k in want_sizes.items():
            assert siz[r] == k, (i, r)
        n_images += 1
    assert n_images == 12000


def test_edge_list_sees_the_colour_agnostic_bridge():
    # the hand-checked bridge of test_oracle_kats: two white pixels three
    # columns apart share a label through block (1, 1)'s empty white set
    t = np.full((8, 8), 127, np.uint8)
    t[1, 1] = t[1, 4] = 255
    t[2, 2] = t[2, 3] = 0
    lab, sizes = bke_reference_labels(t)
    assert lab[1, 1] == lab[1, 4] == 0x80000000 and sizes[0] == 2


# ---------------------------------------------------------------- do_indexing
def _choose(n, k):
    return math.comb(n, k) if n >= k else 0


def textbook_subset(num):
    """The 4-subset {c0 > c1 > c2 > c3} with combination number num =
    C(c0,4) + C(c1,3) + C(c2,2) + C(c3,1)."""
    out, rest = [], num
    for k in (4, 3, 2, 1):
        c = k - 1
        while _choose(c + 1, k) <= rest:
            c += 1
        out.append(c)
        rest -= _choose(c, k)
    return out


def test_textbook_subsets_cover_every_subset_once():
    subs = [tuple(sorted(textbook_subset(n))) for n in range(_choose(10, 4))]
    assert len(set(subs)) == 210 and set(subs) == set(combinations(range(10), 4))


def _moment(oracle, pre, n, i0, i1):
    out = np.zeros(1, R.LINE_FIT_POINT)
    tot = C.c_int64()
    oracle.lib.orc_get_moment(pre.ctypes.data_as(P), n, i0, i1, out.ctypes.data_as(P), C.byref(tot))
    return out[0], tot.value


def _fit(oracle, m, num):
    m = np.array([m], R.LINE_FIT_POINT)
    line = np.zeros(4, np.float32)
    err, mse = C.c_float(), C.c_float()
    oracle.lib.orc_fit_line(m.ctypes.data_as(P), num, line.ctypes.data_as(P), C.byref(err),
                            C.byref(mse))
    return line, np.float32(err.value), np.float32(mse.value)


def _cluster(rng, quiet=False):
    """Angle-ordered boundary points (doubled coordinates, weights) around a
    random noisy quad, as line-fit records, and their cyclic prefix sums
    (quiet: sub-pixel noise, so that many 4-subsets fit within 0.1 of the
    best and the rule's first-within-0.1 choice differs from the argmin)."""
    cx, cy = rng.uniform(200, 800, 2)
    side = rng.uniform(20, 80)
    rot = rng.uniform(0, 2 * np.pi)
    corners = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float) * side / 2
    corners = corners @ np.array([[np.cos(rot), -np.sin(rot)], [np.sin(rot), np.cos(rot)]]).T
    corners += rng.normal(0, side * 0.08, corners.shape) + [cx, cy]
    pts, starts = [], []
    for s in range(4):
        a, b = corners[s], corners[(s + 1) % 4]
        k = int(rng.integers(8, 24))
        starts.append(len(pts))
        for t in np.arange(k) / k:
            sd = 0.05 if quiet else rng.choice([0.3, 1.5, 4.0])
            pts.append(a + t * (b - a) + rng.normal(0, sd, 2))
    pts = np.array(pts)
    rx = np.rint(2 * pts[:, 0]).astype(np.int64)
    ry = np.rint(2 * pts[:, 1]).astype(np.int64)
    ws = rng.integers(1, 6, len(pts)).astype(np.int64)
    raw = np.zeros(len(pts), R.LINE_FIT_POINT)
    raw["Mx"], raw["My"], raw["W"] = ws * rx, ws * ry, ws
    raw["Mxx"], raw["Myy"], raw["Mxy"] = ws * rx * rx, ws * ry * ry, ws * rx * ry
    pre = raw.copy()
    for f in ("Mx", "My", "Mxx", "Myy", "Mxy", "W"):
        pre[f] = np.cumsum(raw[f])
    return pre, starts


def _candidates(rng, n, starts, k):
    """k distinct corner candidates: indices at or next to the true
    corners, then random others."""
    cand = []
    for st in starts:
        for d in rng.permutation([0, 1, -1, 2])[:int(rng.integers(1, 3))]:
            j = (st + int(d)) % n
            if j not in cand:
                cand.append(j)
    while len(cand) < k:
        j = int(rng.integers(n))
        if j not in cand:
            cand.append(j)
    rng.shuffle(cand)
    return cand


def _brute_force(oracle, pre, cand):
    """The reference's choice for one segment: cand = the segment's corner
    point indices in corner-sort order (error descending)."""
    n = len(pre)
    cos_crit = np.float32(math.cos(10 * math.pi / 180))
    m = min(len(cand), 10)
    rows = []
    for num in range(_choose(m, 4)):
        li = sorted(cand[c] for c in textbook_subset(num))
        total, ok, fits = np.float32(0), True, []
        lines = []
        for s in range(4):
            mom, tot = _moment(oracle, pre, n, li[s], li[(s + 1) % 4])
            line, err, mse = _fit(oracle, mom, tot)
            ok &= bool(mse <= np.float32(10))
            total = np.float32(total + err)
            fits.append((mom, tot))
            lines.append(line)
        dot = np.float32(np.float32(lines[0][2] * lines[1][2]) + np.float32(lines[0][3] * lines[1][3]))
        angle_ok = bool(abs(dot) <= cos_crit)
        rows.append((num, total if ok and angle_ok else np.float32(np.inf), li, fits, ok, angle_ok))
    lowest = min(r[1] for r in rows)
    if lowest == np.inf:
        return None, rows
    for r in rows:
        if float(np.float32(r[1] - lowest)) < 0.1:
            return r, rows
    raise AssertionError("no combination within 0.1 of the minimum")


def test_do_indexing_selection_rule_brute_force(oracle):
    rng = np.random.default_rng(77)
    seen = dict(non_argmin=0, mse_reject=0, angle_reject=0, capped=0, none_valid=0, checked=0)
    for trial in range(60):
        pres, cands = [], []
        for c in range(6):
            pre, starts = _cluster(rng, quiet=c % 2 == 1)
            cand = _candidates(rng, len(pre), starts, int(rng.integers(4, 14)))
            pres.append(pre)
            cands.append(cand)
        lfp = np.concatenate(pres)
        cext = np.zeros(len(pres), R.CLUSTER_EXTENTS)
        cext["count"] = [len(p) for p in pres]
        cext["start"] = np.r_[0, np.cumsum(cext["count"])[:-1]]
        corners = np.zeros(sum(len(c) for c in cands), R.CORNER)
        seg = np.zeros(len(pres), R.PEAK_EXTENTS)
        o = 0
        for e, cand in enumerate(cands):
            seg[e] = (o, len(cand))
            for j, idx in enumerate(cand):
                corners[o + j] = (idx, e, np.float32(100 - j))  # error descending
            o += len(cand)
        fitted, _ = oracle.do_indexing(seg, corners, lfp, cext)
        for e, cand in enumerate(cands):
            pick, rows = _brute_force(oracle, pres[e], cand)
            seen["checked"] += 1
            seen["capped"] += len(cand) > 10
            seen["mse_reject"] += any(not r[4] for r in rows)
            seen["angle_reject"] += any(r[4] and not r[5] for r in rows)
            got = fitted[e]
            if pick is None:
                seen["none_valid"] += 1
                assert (got["indices"] == 0).all() and (got["num_in_moments"] == 0).all()
                continue
            argmin = min(rows, key=lambda r: r[1])
            seen["non_argmin"] += argmin[0] != pick[0]
            num, total, li, fits, _, _ = pick
            assert list(got["indices"]) == li, (trial, e)
            for s in range(4):
                assert got["num_in_moments"][s] == fits[s][1]
                assert got["moments"][s].tobytes() == np.array([fits[s][0]]).tobytes()
    # the cases exercise the rejections and the 10-corner cap
    assert seen["mse_reject"] > 0 and seen["angle_reject"] > 0, seen
    assert seen["capped"] > 0 and seen["checked"] == 360, seen


def _square_cluster(weights_alt=1):
    """An axis-aligned square, 16 points a side (doubled coordinates, 2
    apart), every point exactly on its side's line with weight 100, except
    the second point of side 0, weight weights_alt. Returns the prefix
    records and the index of each side's first point (the corners)."""
    x0, y0, step, k = 400, 300, 2, 16
    pts, starts = [], []
    for s_, (dx, dy) in enumerate(((1, 0), (0, 1), (-1, 0), (0, -1))):
        starts.append(len(pts))
        for j in range(k):
            pts.append((x0, y0))
            x0, y0 = x0 + dx * step, y0 + dy * step
    pts = np.array(pts, np.int64)
    ws = np.full(len(pts), 100, np.int64)
    ws[starts[0] + 1] = weights_alt
    raw = np.zeros(len(pts), R.LINE_FIT_POINT)
    rx, ry = pts[:, 0], pts[:, 1]
    raw["Mx"], raw["My"], raw["W"] = ws * rx, ws * ry, ws
    raw["Mxx"], raw["Myy"], raw["Mxy"] = ws * rx * rx, ws * ry * ry, ws * rx * ry
    pre = raw.copy()
    for f in ("Mx", "My", "Mxx", "Myy", "Mxy", "W"):
        pre[f] = np.cumsum(raw[f])
    return pre, starts


def test_do_indexing_takes_first_combination_within_0_1_not_the_argmin(oracle):
    """Corner candidates (corner-sort order) [alt, c1, c2, c3, c0]: the
    true corners {c0..c3} fit with error 0 but have combination number 4;
    {alt, c1, c2, c3} (alt = the point after c0, 1 px off side 3's line,
    weight 1 of 100) has number 0 and an error below 0.1. The reference
    takes number 0."""
    pre, st = _square_cluster()
    cand = [st[0] + 1, st[1], st[2], st[3], st[0]]
    pick, rows = _brute_force(oracle, pre, cand)
    errs = {r[0]: float(r[1]) for r in rows}
    assert errs[4] == 0.0 and 0.0 < errs[0] < 0.1, errs
    assert pick[0] == 0 and min(rows, key=lambda r: r[1])[0] == 4
    cext = np.array([(0, len(pre))], R.CLUSTER_EXTENTS)
    corners = np.array([(c, 0, np.float32(10 - j)) for j, c in enumerate(cand)], R.CORNER)
    seg = np.array([(0, len(cand))], R.PEAK_EXTENTS)
    fitted, _ = oracle.do_indexing(seg, corners, pre, cext)
    assert list(fitted[0]["indices"]) == sorted(cand[:4])


# ---------------------------------------------------------------- process_quads
def _side_moment(a2, b2):
    """Line-fit moment of every lattice point (doubled coordinates) on the
    segment a2 -> b2 (integer endpoints), weight 1: exactly collinear."""
    d = b2 - a2
    g = math.gcd(int(abs(d[0])), int(abs(d[1])))
    pts = np.array([a2 + j * d // g for j in range(g + 1)], np.int64)
    m = np.zeros(1, R.LINE_FIT_POINT)[0]
    m["Mx"], m["My"] = pts[:, 0].sum(), pts[:, 1].sum()
    m["Mxx"], m["Myy"], m["Mxy"] = (pts[:, 0] ** 2).sum(), (pts[:, 1] ** 2).sum(), (pts[:, 0] * pts[:, 1]).sum()
    m["W"] = len(pts)
    return m, len(pts)


def _fitted_from_corners(corners_px):
    """A FittedQuad whose line i is fitted to the points of the side from
    corner i-1 to corner i (process_fitted_quads intersects lines i and
    i+1 into corner i)."""
    q = np.zeros(1, R.FITTED_QUAD)[0]
    c2 = np.rint(np.asarray(corners_px) * 2).astype(np.int64)
    for i in range(4):
        m, n = _side_moment(c2[(i - 1) % 4], c2[i])
        q["moments"][i] = m
        q["num_in_moments"][i] = n
    q["indices"] = (1, 2, 3, 4)
    return q


def _quad(rng, side, rot, centre=(300.0, 200.0), scale=40):
    sq = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float) * side / 2
    R2 = np.array([[math.cos(rot), -math.sin(rot)], [math.sin(rot), math.cos(rot)]])
    c = sq @ R2.T + centre
    # corners on the half-pixel lattice, sides with many lattice points
    return np.rint(c * 2 * scale) / (2 * scale) if scale else c


def _winding_cross(c):
    d1, d2 = c[1] - c[0], c[2] - c[1]
    return d1[0] * d2[1] - d1[1] * d2[0]


def test_process_quads_corners_are_the_side_intersections(oracle):
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(200):
        side = rng.uniform(12, 200)
        c = _quad(rng, side, rng.uniform(0, 2 * math.pi), scale=0)
        c = np.rint(c * 2) / 2 + rng.integers(-3, 4, (4, 2)) / 2.0  # a general convex quad
        if _winding_cross(c) < 0:
            c = c[::-1].copy()
        # convex and well away from the 10 degree and area limits
        angles = []
        for i in range(4):
            u, v = c[(i + 1) % 4] - c[i], c[(i + 2) % 4] - c[(i + 1) % 4]
            angles.append(abs(u @ v) / (np.linalg.norm(u) * np.linalg.norm(v)))
            if u[0] * v[1] - u[1] * v[0] <= 0:
                break
        else:
            if max(angles) > math.cos(math.radians(20)):
                continue
            out = oracle.process_quads(np.array([_fitted_from_corners(c)]))
            assert len(out) == 1
            np.testing.assert_allclose(out["corners"][0], c, atol=1e-3)
            checked += 1
    assert checked > 100


def test_process_quads_rejections(oracle):
    def run(c):
        return len(oracle.process_quads(np.array([_fitted_from_corners(np.asarray(c, float))])))

    sq = [[100, 100], [140, 100], [140, 140], [100, 140]]
    assert _winding_cross(np.array(sq, float)) > 0 and run(sq) == 1
    assert run(sq[::-1]) == 0                       # reversed winding: not convex
    # area: 8 x 8 passes, 7.5 x 7.5 (56.25 < 0.95 * 64) does not
    assert run([[10, 10], [18, 10], [18, 18], [10, 18]]) == 1
    assert run([[10, 10], [17.5, 10], [17.5, 17.5], [10, 17.5]]) == 0
    # angles: a rhombus of 15 degrees passes, one of 8 degrees does not
    for deg, want in ((15, 1), (8, 0)):
        a = math.radians(deg)
        u = np.array([60.0, 0.0])
        v = 60 * np.array([math.cos(a), math.sin(a)])
        c = np.array([[50, 50], [50, 50] + u, [50, 50] + u + v, [50, 50] + v])
        c = np.rint(c * 2) / 2
        assert run(c) == want, deg
    # lines 0 and 1 parallel: corner 0 has no intersection (|det| < 0.001)
    q = _fitted_from_corners(np.array(sq, float))
    q["moments"][1] = _side_moment(np.array([200, 200]), np.array([280, 200]))[0]  # parallel to side 0
    q["moments"][0] = _side_moment(np.array([200, 300]), np.array([280, 300]))[0]
    assert len(oracle.process_quads(np.array([q]))) == 0
    # never-written fits (all indices 0) are skipped
    q = _fitted_from_corners(np.array(sq, float))
    q["indices"] = 0
    assert len(oracle.process_quads(np.array([q]))) == 0


# ---------------------------------------------------------------- float64 twin
def test_float32_corners_within_1e3_px_of_the_float64_twin(oracle):
    import pyoracle
    from pyatq import synth
    for fr, f in ((synth.plain_small(), 2), (synth.plain_1080p(2000), 1), (synth.plain_1080p(2001), 2)):
        out = oracle.pipeline(fr.image, f)
        err, n = pyoracle.corner_error_fp64(out["fitted"], out["quads"])
        assert n > 0 and err <= 1e-3, (err, n)


def test_float64_twin_is_exact_on_collinear_sides():
    import pyoracle
    c = np.array([[100.5, 100.0], [150.0, 104.5], [146.0, 152.0], [98.0, 147.5]])
    got = pyoracle.quad_corners_fp64(_fitted_from_corners(c))
    np.testing.assert_allclose(got, c, atol=1e-9)


def test_oracle_widens_cluster_ids_past_u16(oracle):
    """73,790 valid clusters (one per dark square): the reference's u16
    cluster index (65535 = sentinel, src/main.cpp:639-651) cannot name them;
    the oracle's u32 ids give every square its own cluster and quad."""
    from pyatq import synth
    img = synth.square_grid().image
    q, st = oracle.detect(img, 1, max_out=1 << 17)
    assert st["n_clusters"] == 73790
    assert len(q) > 65535
    assert int(q["blob_index"].max()) > 65535
    assert len(np.unique(q["blob_index"])) == len(q)
