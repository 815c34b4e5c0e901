"""The oracle against the reference's own code: golden vectors written from
the reference's SYCL-free headers (tests/golden/make_ref_vectors.py) and, when
the reference checkout has been compiled into oracle/_ref, the live functions.

Pins calc_theta, the point field accessors (the zero half-pixel mask),
ClusterBounds inital_from_point / reduce_bounds / cx / cy / dot,
ValidBlobFilter, compute_initial_linefit (its row index taken from rx),
Combinadics<uint8_t, 15, 4, 16> (n_choose_k, decode), and the loader's
colour-to-grey conversion (stb_image's, tests/golden/make_grey_vectors.py).
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN

P = C.c_void_p
V = np.load(os.path.join(GOLDEN, "ref_header_vectors.npz"))
GV = np.load(os.path.join(GOLDEN, "ref_grey_vectors.npz"))
REF_STB = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libref_stb.so")


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_record_sizes_match_reference():
    from pyatq import records as R
    # BoundaryPoint, ClusterPoint, ClusterBounds, LineFitPoint, ClusterExtents
    assert list(V["ref_sizeof"]) == [R.BOUNDARY_POINT.itemsize, R.CLUSTER_POINT.itemsize,
                                     R.CLUSTER_BOUNDS.itemsize, R.LINE_FIT_POINT.itemsize,
                                     R.CLUSTER_EXTENTS.itemsize]


def test_calc_theta_bit_exact(oracle):
    got = np.array([oracle.lib.orc_calc_theta(int(a), int(b), float(c), float(d)) for a, b, c, d in
                    zip(V["theta_px"], V["theta_py"], V["theta_cx"], V["theta_cy"])], np.float32)
    np.testing.assert_array_equal(_bits(got), _bits(V["theta"]))


def test_bounds_from_point_and_reduce(oracle):
    from pyatq import records as R
    px, py, st = V["bp_px"], V["bp_py"], V["bp_start"]
    b1 = np.zeros(len(px), R.CLUSTER_BOUNDS)
    red = np.zeros(len(px), R.CLUSTER_BOUNDS)
    b2 = V["bounds_b2"]
    for i in range(len(px)):
        oracle.lib.orc_bounds_from_point(int(px[i]), int(py[i]), int(st[i]), b1[i:i + 1].ctypes.data_as(P))
        oracle.lib.orc_reduce_bounds(b1[i:i + 1].ctypes.data_as(P), b2[i:i + 1].ctypes.data_as(P),
                                     red[i:i + 1].ctypes.data_as(P))
    assert b1.tobytes() == V["bounds_from_point"].tobytes()
    assert red.tobytes() == V["bounds_reduced"].tobytes()


def test_point_fields_zero_half_pixel_mask():
    # the reference's half-pixel mask is (1<<15)&(1<<14) == 0: every record is
    # read back as RIGHT, so x_with_dx = 2x+1, y_with_dy = 2y, dx = 1, dy = 0
    f = V["point_fields"]
    x = (V["theta_px"] & 0x3FFF).astype(np.int64)
    y = (V["theta_py"] & 0x3FFF).astype(np.int64)
    b2w = (V["theta_py"] >> 15).astype(np.int64)
    np.testing.assert_array_equal(f[:, 0], 2 * x + 1)
    np.testing.assert_array_equal(f[:, 1], 2 * y)
    np.testing.assert_array_equal(f[:, 2], 1)
    np.testing.assert_array_equal(f[:, 3], 0)
    np.testing.assert_array_equal(f[:, 4], np.where(b2w == 1, 1, -1))
    np.testing.assert_array_equal(f[:, 5], 0)


def test_cx_cy_dot_and_filter(oracle):
    rb = V["rand_bounds"]
    got = np.zeros((len(rb), 3), np.float32)
    for i in range(len(rb)):
        oracle.lib.orc_bounds_cx_cy_dot(rb[i:i + 1].ctypes.data_as(P), got[i].ctypes.data_as(P))
    np.testing.assert_array_equal(_bits(got), _bits(V["cx_cy_dot"]))
    for k, (w, h) in enumerate(V["valid_dims"]):
        ok = np.array([oracle.lib.orc_valid_blob(rb[i:i + 1].ctypes.data_as(P), int(w), int(h))
                       for i in range(len(rb))])
        np.testing.assert_array_equal(ok, V["valid_blob"][:, k])
    assert V["valid_blob"].sum() > 100  # the vectors exercise both outcomes
    assert (V["valid_blob"] == 0).sum() > 100


def test_initial_linefit(oracle):
    from pyatq import records as R
    g = np.ascontiguousarray(V["lf_grey"])
    gh, gw = g.shape
    out = np.zeros(len(V["lf_px"]), R.LINE_FIT_POINT)
    for i, (a, b) in enumerate(zip(V["lf_px"], V["lf_py"])):
        oracle.lib.orc_initial_linefit(int(a), int(b), gw, gh, g.ctypes.data_as(P),
                                       out[i:i + 1].ctypes.data_as(P))
    assert out.tobytes() == V["linefit"].tobytes()


def test_combinadics(oracle):
    for num in range(256):
        o = np.zeros(4, np.int64)
        oracle.lib.orc_combinadics_decode(num, o.ctypes.data_as(P))
        np.testing.assert_array_equal(o, V["comb_decode"][num])
    nck = np.array([[oracle.lib.orc_combinadics_n_choose_k(n, k) for n in range(16)] for k in range(5)])
    np.testing.assert_array_equal(nck, V["comb_nck"])


def test_combinadics_closed_form():
    """For every combination number below C(10,4) the uint8 table decodes to
    the textbook combinatorial number system (c0 > c1 > c2 > c3)."""
    from math import comb
    for num in range(210):
        n, cf = num, []
        for kk in (4, 3, 2, 1):
            c = kk - 1
            while comb(c + 1, kk) <= n:
                c += 1
            cf.append(c)
            n -= comb(c, kk)
        assert list(V["comb_decode"][num]) == cf
    assert [int(V["comb_nck"][4][m]) for m in range(11)] == [comb(m, 4) for m in range(11)]


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref",
                                                    "libref_headers.so")),
                    reason="reference headers not compiled here (oracle/_ref)")
def test_live_reference_headers_agree_on_fresh_inputs(oracle):
    import pyoracle
    ref = pyoracle.RefHeaders().lib
    rng = np.random.default_rng(99)
    for _ in range(2000):
        px = int(rng.integers(0, 16384)) | (int(rng.integers(0, 4)) << 14)
        py = int(rng.integers(0, 16384)) | (int(rng.integers(0, 2)) << 15)
        cx, cy = float(rng.uniform(-10, 40000)), float(rng.uniform(-10, 40000))
        a = np.float32(ref.ref_calc_theta(px, py, cx, cy))
        b = np.float32(oracle.lib.orc_calc_theta(px, py, cx, cy))
        assert a.view(np.uint32) == b.view(np.uint32)


def test_grey_conversion_matches_reference_loader():
    import pyoracle
    for px, want in ((GV["rgb"], GV["rgb_grey"]), (GV["rgba"], GV["rgba_grey"])):
        got = pyoracle.grey_from_pixels(px[None, :, :])[0]
        np.testing.assert_array_equal(got, want)
    # closed-form corners: black, white, the primaries
    assert list(pyoracle.grey_from_pixels(np.array([[[0, 0, 0], [255, 255, 255], [255, 0, 0],
                                                     [0, 255, 0], [0, 0, 255]]], np.uint8))[0]) == \
        [0, 255, 76, 149, 28]


@pytest.mark.skipif(not os.path.exists(REF_STB), reason="reference loader not compiled here (oracle/_ref)")
def test_live_reference_loader_agrees_on_fresh_pixels():
    import pyoracle
    ref = pyoracle.RefStb()
    rng = np.random.default_rng(7)
    for c in (3, 4):
        px = rng.integers(0, 256, (20000, c), dtype=np.uint8)
        np.testing.assert_array_equal(pyoracle.grey_from_pixels(px[None])[0], ref.grey(px, c))
