"""Whole-pipeline parity on the GPU: the detector's every intermediate and its
quads against the oracle (grouping through process_fitted_quads), and the
glue drop-ins chained.

Bit-exact everywhere, element by element: the GPU writes boundary records
in the reference's plane-major raster order and every sort is stable, as the
oracle's are. The detector drops, before its key sort, the records of pairs
that ValidBlobFilter rejects; its grouped records are compared with the
oracle's restricted to the valid clusters.
"""
import numpy as np
import pytest

from conftest import REF_PNGS, load_png_gray

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def atq():
    import pyatq
    pyatq.init()
    return pyatq


def _frames():
    from pyatq import synth
    out = [("plain_small_f2", synth.plain_small().image, 2),
           ("plain_1080p_f1", synth.plain_1080p(2000).image, 1),
           ("plain_1080p_f2", synth.plain_1080p(2001).image, 2),
           ("plain_1080p_f1_b", synth.plain_1080p(2003).image, 1)]
    for name in REF_PNGS:
        out.append((name, load_png_gray(name), 1))
    return out


FRAMES = _frames()


def run_detector(atq, img, f, **kw):
    from pyatq import records as R
    h_in, w_in = img.shape
    det = atq.Detector(w_in, h_in, decimate=f, **kw)
    dev = atq.dev_from_np(img)
    det.detect(dev, w_in, h_in)
    atq.sync()
    c = det.counts()
    out = {"counts": c, "dims": det.dims()}
    rd = det.read
    npx = det.dims()[0] * det.dims()[1]
    out["thresholded"] = rd("THRESHOLDED", np.uint8, npx)
    out["labels"] = rd("LABELS", np.uint32, npx)
    out["sizes"] = rd("SIZES", np.uint32, npx)
    nb = int(c["n_boundary"])
    out["points"] = rd("BOUNDARY_POINTS", np.uint32, nb)
    packed = rd("BOUNDARY_KEYS", np.uint64, nb) >> np.uint64(2)  # low 2 bits: plane
    lb = np.uint64(int(c["key_label_bits"]))
    out["keys"] = ((packed >> lb) << np.uint64(32)) | (packed & ((np.uint64(1) << lb) - np.uint64(1)))
    ncl = int(c["n_clusters"])
    out["bounds"] = rd("CLUSTER_BOUNDS", R.CLUSTER_BOUNDS, ncl)
    out["extents"] = rd("CLUSTER_EXTENTS", R.CLUSTER_EXTENTS, ncl)
    nfp = int(c["n_cluster_points"])
    out["cluster_points"] = rd("CLUSTER_POINTS", R.CLUSTER_POINT, nfp)
    out["cluster_index"] = rd("CLUSTER_INDEX", np.uint32, nfp)
    out["line_fit"] = rd("LINE_FIT", R.LINE_FIT_POINT, nfp)
    out["sparse_corners"] = rd("SPARSE_CORNERS", R.CORNER, nfp)
    out["corners"] = rd("CORNERS", R.CORNER, int(c["n_corners"]))
    ne = int(c["n_segments"])
    out["peak_extents"] = rd("PEAK_EXTENTS", R.PEAK_EXTENTS, ne)
    out["fitted"] = rd("FITTED", R.FITTED_QUAD, ne)
    out["quads"] = rd("QUADS", R.QUAD_CORNERS, int(c["n_quads"]))
    slots = det.buffer("PAIR_KEYS")[1]
    out["pair_keys"] = rd("PAIR_KEYS", np.uint64, slots)
    out["pair_counts"] = rd("PAIR_COUNTS", np.uint32, slots)
    det.close()
    return out


def _bytes(a):
    return np.ascontiguousarray(a).view(np.uint8)


def surviving_records(o):
    """The oracle's grouped records restricted to the pairs ValidBlobFilter
    accepts (the oracle's cluster bounds): what the detector keeps when its
    pair table did not overflow. Also the position of every oracle record in
    that restricted array."""
    keys, points = o["keys"], o["points"]
    keep = np.zeros(len(keys), bool)
    for b in o["bounds"]:
        keep[int(b["start"]):int(b["start"]) + int(b["count"])] = True
    new_pos = np.cumsum(keep) - keep
    return keys[keep], points[keep], len(o["bounds"]), new_pos


def check_pair_table(g, o):
    """Every pair of the frame in the table, with its exact record count up
    to the rejection limit and a count above it otherwise."""
    w, h = g["dims"]
    limit = 2 * (2 * w + 2 * h)
    uniq, cnt = np.unique(o["keys"], return_counts=True)
    used = g["pair_keys"] != 0          # slots hold ~key
    keys = ~g["pair_keys"][used]
    order = np.argsort(keys)
    gk, gc = keys[order], g["pair_counts"][used][order]
    np.testing.assert_array_equal(gk, uniq)
    np.testing.assert_array_equal(np.minimum(gc, limit + 1), np.minimum(cnt, limit + 1))


def expected_sort_passes(n_valid, slots, filter_off):
    """Radix passes over the pair-rank keys (11-bit digits over the rank's
    bits, plus the plane's 2 bits unless dropping them saves a pass; at most
    2): 0 selects the full-key sort, taken when the pair table overflowed or
    more pairs are valid than the detector ranks (a quarter of the table's
    slots, 64 to 16384)."""
    cap = max(64, min(16384, slots // 4))
    if filter_off or n_valid > cap:
        return 0
    bits = max(n_valid - 1, 0).bit_length()
    return max(1, -(-bits // 11))


def check_labels_and_sizes(g, o, min_size=25):
    """Labels bit-exact; the detector's sizes are exact below min_size and
    only "at least min_size" above it (its size adds stop there)."""
    np.testing.assert_array_equal(g["labels"], o["labels"].reshape(-1), err_msg="labels")
    roots = np.unique(np.concatenate([[0], o["labels"].reshape(-1) & 0x7FFFFFFF]))
    want = o["sizes"].reshape(-1)[roots]
    got = g["sizes"][roots]
    small = want < min_size
    np.testing.assert_array_equal(got[small], want[small], err_msg="sizes below min_size")
    assert (got[~small] >= min_size).all() and (got[~small] <= want[~small]).all()


def check_detector_corners(got, want, peaks):
    """The detector orders only what do_indexing reads: each segment's
    min(10, n) first peaks equal the reference's sorted ones exactly; the
    segment's other peaks are the reference's others, in some order."""
    assert len(got) == len(want)
    starts = peaks["start"].astype(np.int64)
    counts = peaks["count"].astype(np.int64)
    assert counts.sum() == len(want)
    seg = np.repeat(np.arange(len(counts)), counts)
    within = np.arange(len(want)) - np.repeat(starts, counts)
    top = within < 10
    np.testing.assert_array_equal(_bytes(got[top]), _bytes(want[top]), err_msg="top peaks")
    rest = np.flatnonzero(~top)
    if len(rest):
        a, b = got[rest], want[rest]
        oa = np.lexsort((a["line_fit_point_index"], seg[rest]))
        ob = np.lexsort((b["line_fit_point_index"], seg[rest]))
        np.testing.assert_array_equal(_bytes(a[oa]), _bytes(b[ob]), err_msg="other peaks")


def check_against_oracle(g, o):
    c = g["counts"]
    assert int(c["overflow"]) == 0
    w, h = g["dims"]
    filter_off = int(c["pair_filter_off"])
    if not filter_off:
        check_pair_table(g, o)
        assert int(c["n_valid_pairs"]) == len(o["bounds"])
    passes = expected_sort_passes(len(o["bounds"]), len(g["pair_keys"]), filter_off)
    assert int(c["sort_passes"]) == passes
    if passes:
        # the plane leaves the key (records arrive plane by plane) only when
        # that saves a pass: 513 to 2048 valid pairs
        n = len(o["bounds"])
        assert int(c["rank_shift"]) == (0 if 512 < n <= 2048 else 2)
    o_bounds = o["bounds"]
    if int(c["pair_filter_off"]):
        keys, points, n_pairs = o["keys"], o["points"], len(o["bounds_all"])
    else:
        keys, points, n_pairs, new_pos = surviving_records(o)
        # a cluster's start indexes the detector's own (restricted) records
        o_bounds = o_bounds.copy()
        if len(o_bounds):
            o_bounds["start"] = new_pos[o_bounds["start"]]
    assert int(c["n_boundary"]) == len(points)
    assert int(c["n_pairs"]) == n_pairs
    assert int(c["n_clusters"]) == len(o["bounds"])
    assert int(c["n_cluster_points"]) == len(o["cluster_points"])
    assert int(c["n_corners"]) == len(o["corners"])
    assert int(c["n_segments"]) == len(o["peak_extents"])
    assert int(c["oob_reads"]) == int(o["oob_reads"])
    np.testing.assert_array_equal(g["thresholded"], o["thresholded"].reshape(-1))
    check_labels_and_sizes(g, o)
    # grouped records: the reference's stable order, element by element
    np.testing.assert_array_equal(g["keys"], keys)
    np.testing.assert_array_equal(g["points"], points)
    np.testing.assert_array_equal(_bytes(g["bounds"]), _bytes(o_bounds), err_msg="bounds")
    for name in ("extents", "cluster_points", "cluster_index", "line_fit",
                 "sparse_corners", "peak_extents", "fitted"):
        np.testing.assert_array_equal(_bytes(g[name]), _bytes(o[name]), err_msg=name)
    check_detector_corners(g["corners"], o["corners"], o["peak_extents"])
    assert int(c["n_quads"]) == len(o["quads"])
    np.testing.assert_array_equal(_bytes(g["quads"]), _bytes(o["quads"]))


@pytest.mark.parametrize("name,img,f", FRAMES, ids=[x[0] for x in FRAMES])
def test_detector_matches_oracle(atq, oracle, name, img, f):
    o = oracle.pipeline(img, f)
    g = run_detector(atq, img, f)
    check_against_oracle(g, o)
    q, _ = oracle.detect(img, f)
    np.testing.assert_array_equal(_bytes(g["quads"]), _bytes(q))


def test_detector_reuse_shrinking_frames(atq, oracle):
    """One detector, frames of decreasing size: no state leaks between frames."""
    from pyatq import synth
    big = synth.plain_1080p(2004).image
    small = synth.plain_small(1001).image
    det = atq.Detector(1920, 1080, decimate=1)
    for img in (big, small[:240, :320].copy(), big[:480, :640].copy()):
        h, w = img.shape
        det.detect(atq.dev_from_np(img), w, h)
        atq.sync()
        got = det.quads()
        want, _ = oracle.detect(img, 1)
        np.testing.assert_array_equal(_bytes(got), _bytes(want))
    det.close()


def test_detector_pair_table_overflow(atq, oracle):
    """5x5 squares on a 2-pixel grid make thousands of cluster pairs; a
    detector with a 256-slot pair table must notice the overflow, keep every
    boundary record, and still match the oracle element by element."""
    yy, xx = np.mgrid[0:480, 0:640]
    img = np.where(((yy % 7) < 5) & ((xx % 7) < 5), 220, 30).astype(np.uint8)
    o = oracle.pipeline(img, 1)
    assert len(np.unique(o["keys"])) > 1000
    g = run_detector(atq, img, 1, pair_table_log2=8)
    assert int(g["counts"]["pair_filter_off"]) == 1
    assert int(g["counts"]["sort_passes"]) == 0  # the full-key sort
    check_against_oracle(g, o)
    # the default table holds them all: filtered, and the same downstream
    g = run_detector(atq, img, 1)
    assert int(g["counts"]["pair_filter_off"]) == 0
    check_against_oracle(g, o)


def test_detector_more_valid_pairs_than_ranked(atq, oracle):
    """A 512-slot pair table holds the 286 pairs of a 1080p frame without
    overflow but ranks at most 128 pairs: with 140 valid pairs the records
    keep their packed pair keys, the full-key sort runs, and every
    intermediate still matches the oracle."""
    from pyatq import synth
    img = synth.plain_1080p(2000).image
    o = oracle.pipeline(img, 1)
    assert len(o["bounds"]) > 128
    g = run_detector(atq, img, 1, pair_table_log2=9)
    assert int(g["counts"]["pair_filter_off"]) == 0
    assert int(g["counts"]["sort_passes"]) == 0
    check_against_oracle(g, o)


DEGENERATE = [
    ("flat_64x48", np.full((48, 64), 100, np.uint8)),
    ("noise_8x8", np.random.default_rng(7).integers(0, 256, (8, 8), dtype=np.uint8)),
    ("one_square_32x32", np.pad(np.full((12, 12), 30, np.uint8), 10, constant_values=220)),
]


@pytest.mark.parametrize("name,img", DEGENERATE, ids=[x[0] for x in DEGENERATE])
def test_detector_degenerate_frames(atq, oracle, name, img):
    """Frames with no or few cluster pairs: the ranking, the one-pass sort
    of zero or a handful of records and every later stage still agree with
    the oracle element by element."""
    o = oracle.pipeline(img, 1)
    g = run_detector(atq, img, 1)
    check_against_oracle(g, o)


def test_detector_stress_12mp(atq, oracle):
    """The 4096x3072 stress frame (maze, blobs, 200 tags), element by element:
    its longest clusters exceed one work-group's LDS merge (16384 points), so
    the global merge passes of the segment sort run too."""
    from pyatq import synth
    img = synth.stress_12mp().image
    o = oracle.pipeline(img, 1)
    assert int(o["extents"]["count"].max()) > 16384
    g = run_detector(atq, img, 1)
    check_against_oracle(g, o)  # labels of the maze included
    # the corners' error against a float64 twin of the quad stage
    import pyoracle
    err, n = pyoracle.corner_error_fp64(g["fitted"], g["quads"])
    print(f"12 MP: {n} quads, max |float32 - float64| corner {err:.2e} px")
    assert n > 100 and err < 2e-3


def test_detector_more_valid_clusters_than_u16_ids(atq, oracle):
    """73,790 valid clusters at 4096x3072: more than the reference's u16
    cluster ids (65535 is its sentinel) hold, so every cluster index above
    65,535 goes through the widened u32 ids, element by element against the
    oracle, labels included."""
    from pyatq import synth
    img = synth.square_grid().image
    o = oracle.pipeline(img, 1)
    assert len(o["extents"]) > 65535 and len(o["quads"]) > 65535
    assert int(o["quads"]["blob_index"].max()) > 65535
    g = run_detector(atq, img, 1)
    assert int(g["counts"]["overflow"]) == 0
    check_against_oracle(g, o)


def test_detector_two_rank_passes(atq, oracle):
    """About 7,000 valid clusters at 1280x960: more than one 11-bit radix pass
    of the pair ranks covers (2048), fewer than the detector ranks at this
    table size (8192), so the rank sort takes its second pass."""
    from pyatq import synth
    img = synth.square_grid(1280, 960).image
    o = oracle.pipeline(img, 1)
    assert 2048 < len(o["bounds"]) <= 8192
    g = run_detector(atq, img, 1)
    assert int(g["counts"]["sort_passes"]) == 2
    check_against_oracle(g, o)


def test_every_front_shape_gives_identical_results(atq, oracle):
    """Threshold regions and CCL tiles only split the work differently: the
    thresholded image, the labels and the quads are the same for every shape
    the library takes, and equal to the oracle's."""
    from pyatq import synth
    img = synth.plain_1080p(2005).image
    h_in, w_in = img.shape
    want, _ = oracle.detect(img, 1)
    o_labels, _ = oracle.segmentation(oracle.threshold(img)[2])
    dev = atq.dev_from_np(img)
    det = atq.Detector(w_in, h_in)
    with pytest.raises(atq.AtqError):
        det.set_shapes((48, 8, 16))
    for sh in atq.SHAPES:
        det.set_shapes(sh)
        assert det.shapes() == sh
        det.detect(dev, w_in, h_in)
        atq.sync()
        labels = det.read("LABELS", np.uint32, w_in * h_in)
        np.testing.assert_array_equal(labels, o_labels.reshape(-1), err_msg=str(sh))
        np.testing.assert_array_equal(_bytes(det.quads()), _bytes(want), err_msg=str(sh))
    det.close()


def test_autotune_picks_a_supported_shape(atq, oracle):
    from pyatq import synth
    img = synth.plain_1080p(2006, decimate=2).image
    h_in, w_in = img.shape
    dev = atq.dev_from_np(img)
    det = atq.Detector(w_in, h_in, decimate=2)
    sh = det.autotune(dev, w_in, h_in)
    assert sh in atq.SHAPES and det.shapes() == sh
    det.detect(dev, w_in, h_in)
    atq.sync()
    want, _ = oracle.detect(img, 2)
    np.testing.assert_array_equal(_bytes(det.quads()), _bytes(want))
    det.close()


def _colour(img, channels, seed):
    """An RGB(A) frame whose loader grey (stb_image's) is close to img: the
    grey value spread over the channels with seeded jitter."""
    rng = np.random.default_rng(seed)
    h, w = img.shape
    px = np.repeat(img[:, :, None], channels, axis=2).astype(np.int32)
    px[:, :, :3] += rng.integers(-20, 21, (h, w, 3))
    return np.clip(px, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("channels,f,w_in", [(3, 1, 800), (4, 2, 800), (3, 3, 790)])
def test_colour_front_end_matches_oracle(atq, oracle, channels, f, w_in):
    """RGB(A) input: atq_grey_decimate against the reference loader's
    conversion (the oracle's, pinned to stb_image), and the detector's fused
    colour path against the oracle run on that grey image."""
    import pyoracle
    from pyatq import synth
    base = synth.plain_1080p(2007 + f).image[:600, :w_in].copy()
    px = _colour(base, channels, 11 + f)
    h_in, w_in = base.shape
    grey = pyoracle.grey_from_pixels(px)
    want_grey = oracle.decimate(grey, f) if f > 1 else grey
    dev = atq.dev_from_np(px)
    (w, h), g_dev = atq.grey_decimate(dev, w_in, h_in, channels, f)
    np.testing.assert_array_equal(atq.to_np(g_dev, np.uint8, w * h).reshape(h, w), want_grey)
    det = atq.Detector(w_in, h_in, decimate=f)
    det.detect_rgb(dev, w_in, h_in, channels)
    atq.sync()
    np.testing.assert_array_equal(det.read("GRAY", np.uint8, w * h), want_grey.reshape(-1))
    want, _ = oracle.detect(grey, f)
    np.testing.assert_array_equal(_bytes(det.quads()), _bytes(want))
    det.close()
