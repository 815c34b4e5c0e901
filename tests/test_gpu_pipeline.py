"""Whole-pipeline parity on the GPU: the detector's every intermediate and its
quads against the oracle (rows A8-A21), and the glue drop-ins chained.

Bit-exact everywhere, element by element: the GPU writes boundary records
in the reference's plane-major raster order and every sort is stable, as the
oracle's are.
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


def run_detector(atq, img, f):
    from pyatq import records as R
    h_in, w_in = img.shape
    det = atq.Detector(w_in, h_in, decimate=f)
    dev = atq.dev_from_np(img)
    det.detect(dev, w_in, h_in)
    atq.sync()
    c = det.counts()
    out = {"counts": c}
    rd = det.read
    out["thresholded"] = rd("THRESHOLDED", np.uint8, det.dims()[0] * det.dims()[1])
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
    det.close()
    return out


def _bytes(a):
    return np.ascontiguousarray(a).view(np.uint8)


def check_against_oracle(g, o):
    c = g["counts"]
    assert int(c["overflow"]) == 0
    assert int(c["n_boundary"]) == len(o["points"])
    assert int(c["n_pairs"]) == len(o["bounds_all"])
    assert int(c["n_clusters"]) == len(o["bounds"])
    assert int(c["n_cluster_points"]) == len(o["cluster_points"])
    assert int(c["n_corners"]) == len(o["corners"])
    assert int(c["n_segments"]) == len(o["peak_extents"])
    assert int(c["oob_reads"]) == int(o["oob_reads"])
    np.testing.assert_array_equal(g["thresholded"], o["thresholded"].reshape(-1))
    # grouped records: the reference's stable order, element by element
    np.testing.assert_array_equal(g["keys"], o["keys"])
    np.testing.assert_array_equal(g["points"], o["points"])
    for name in ("bounds", "extents", "cluster_points", "cluster_index", "line_fit",
                 "sparse_corners", "corners", "peak_extents", "fitted"):
        np.testing.assert_array_equal(_bytes(g[name]), _bytes(o[name]), err_msg=name)
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
