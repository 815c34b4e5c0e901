"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product path never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "sycl-apriltag-test_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)
from pyatq import records as R  # noqa: E402

_P = C.c_void_p
_I = C.c_int64


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "width", "height", "n_sparse_points", "n_pairs", "n_clusters", "n_cluster_points",
        "n_corners", "n_segments", "n_quads", "oob_reads", "n_fitted")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def _ptr(a):
    return a.ctypes.data_as(_P)


# Stages with a fixed-mode variant (orc_fx_*): see the fixed-mode section of
# atq_oracle.cpp for what each corrects.
_FIXED_STAGES = ("find_boundaries", "compact_sort", "cluster_bounds", "cluster_points",
                 "linefit", "fit_lines", "do_indexing", "detect")


class Oracle:
    """The oracle's stages; `fixed=True` selects fixed mode (the reference's
    behaviours corrected), the default compat mode restates the reference."""

    def __init__(self, path=None, fixed=False):
        self.fixed = bool(fixed)
        path = path or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        sig = {
            "orc_decimate": (None, [_P, _I, _I, C.c_int, _P]),
            "orc_threshold": (None, [_P, _I, _I, _P, _P, _P]),
            "orc_segmentation": (None, [_P, _I, _I, _P, _P, _P]),
            "orc_find_boundaries": (None, [_P, _P, _I, _I, _P, _P]),
            "orc_compact_sort": (_I, [_P, _P, _I, _P, _P]),
            "orc_cluster_bounds": (_I, [_P, _P, _I, _P]),
            "orc_filter_bounds": (_I, [_P, _I, _I, _I, _P]),
            "orc_cluster_points": (_I, [_P, _I, _P, _I, _P, _P]),
            "orc_cluster_extents": (None, [_P, _I, _P]),
            "orc_angle_sort": (None, [_P, _P, _I]),
            "orc_linefit": (None, [_P, _I, _P, _I, _I, _P]),
            "orc_linefit_prefix": (None, [_P, _P, _I, _P]),
            "orc_fit_lines": (None, [_P, _P, _P, _I, _P]),
            "orc_corner_segments": (_I, [_P, _I, _P, _P, C.POINTER(C.c_int64)]),
            "orc_do_indexing": (_I, [_P, _I, _P, _P, _I, _P, _P]),
            "orc_process_quads": (_I, [_P, _I, _P]),
            "orc_detect": (_I, [_P, _I, _I, C.c_int, C.c_int, _P, _I, C.POINTER(Stats)]),
            "orc_fit_line": (None, [_P, _I, _P, _P, _P]),
            "orc_get_moment": (None, [_P, _I, _I, _I, _P, C.POINTER(C.c_int64)]),
            "orc_calc_theta": (C.c_float, [C.c_uint16, C.c_uint16, C.c_float, C.c_float]),
            "orc_bounds_cx_cy_dot": (None, [_P, _P]),
            "orc_valid_blob": (C.c_int, [_P, _I, _I]),
            "orc_bounds_from_point": (None, [C.c_uint16, C.c_uint16, C.c_uint32, _P]),
            "orc_reduce_bounds": (None, [_P, _P, _P]),
            "orc_initial_linefit": (None, [C.c_uint16, C.c_uint16, _I, _I, _P, _P]),
            "orc_combinadics_n_choose_k": (C.c_int, [C.c_int, C.c_int]),
            "orc_combinadics_decode": (None, [C.c_int, _P]),
            "orc_hypotf": (C.c_float, [C.c_float, C.c_float]),
            "orc_max_threads": (C.c_int, []),
            "orc_fx_segmentation": (None, [_P, _I, _I, _P, _P]),
            "orc_fx_calc_theta": (C.c_float, [C.c_uint16, C.c_uint16, C.c_float, C.c_float]),
            "orc_fx_bounds_from_point": (None, [C.c_uint16, C.c_uint16, C.c_uint32, _P]),
            "orc_fx_initial_linefit": (None, [C.c_uint16, C.c_uint16, _I, _I, _P, _P]),
        }
        for stage in _FIXED_STAGES:
            sig["orc_fx_" + stage] = sig["orc_" + stage]
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    def _f(self, stage):
        prefix = "orc_fx_" if self.fixed and stage in _FIXED_STAGES else "orc_"
        return getattr(self.lib, prefix + stage)

    # ---- stages ------------------------------------------------------------
    def decimate(self, img, f):
        h_in, w_in = img.shape
        w, h = 1 + (w_in - 1) // f, 1 + (h_in - 1) // f
        out = np.empty((h, w), np.uint8)
        img = np.ascontiguousarray(img)
        self.lib.orc_decimate(_ptr(img), w_in, h_in, f, _ptr(out))
        return out

    def threshold(self, gray):
        h, w = gray.shape
        gray = np.ascontiguousarray(gray)
        ext = np.empty((h // 4, w // 4), R.EXTREMA)
        fil = np.empty((h // 4, w // 4), R.EXTREMA)
        thr = np.empty((h, w), np.uint8)
        self.lib.orc_threshold(_ptr(gray), w, h, _ptr(ext), _ptr(fil), _ptr(thr))
        return ext, fil, thr

    def segmentation(self, thr):
        h, w = thr.shape
        thr = np.ascontiguousarray(thr)
        labels = np.empty((h, w), np.uint32)
        if self.fixed:  # sizes[0] = 0; component sizes at root + 1
            sizes = np.empty(h * w + 1, np.uint32)
            self.lib.orc_fx_segmentation(_ptr(thr), w, h, _ptr(labels), _ptr(sizes))
            return labels, sizes
        scratch = np.empty((h, w), np.uint32)
        sizes = np.empty(h * w, np.uint32)
        self.lib.orc_segmentation(_ptr(thr), w, h, _ptr(scratch), _ptr(labels), _ptr(sizes))
        return labels, sizes

    def find_boundaries(self, labels, sizes):
        h, w = labels.shape
        pts = np.empty(4 * w * h, np.uint32)
        keys = np.empty(4 * w * h, np.uint64)
        self._f("find_boundaries")(_ptr(np.ascontiguousarray(labels)), _ptr(sizes), w, h,
                                     _ptr(pts), _ptr(keys))
        return pts, keys

    def compact_sort(self, pts, keys):
        n = len(pts)
        op = np.empty(n, np.uint32)
        ok = np.empty(n, np.uint64)
        m = self._f("compact_sort")(_ptr(pts), _ptr(keys), n, _ptr(op), _ptr(ok))
        return op[:m].copy(), ok[:m].copy()

    def cluster_bounds(self, pts, keys):
        out = np.empty(max(len(pts), 1), R.CLUSTER_BOUNDS)
        m = self._f("cluster_bounds")(_ptr(pts), _ptr(keys), len(pts), _ptr(out))
        return out[:m].copy()

    def filter_bounds(self, bounds, w, h):
        out = np.empty(max(len(bounds), 1), R.CLUSTER_BOUNDS)
        m = self.lib.orc_filter_bounds(_ptr(bounds), len(bounds), w, h, _ptr(out))
        return out[:m].copy()

    def cluster_points(self, pts, filt):
        n = len(pts)
        op = np.empty(max(n, 1), R.CLUSTER_POINT)
        oi = np.empty(max(n, 1), np.uint32)
        m = self._f("cluster_points")(_ptr(pts), n, _ptr(filt), len(filt), _ptr(op), _ptr(oi))
        return op[:m].copy(), oi[:m].copy()

    def cluster_extents(self, filt):
        out = np.empty(max(len(filt), 1), R.CLUSTER_EXTENTS)
        self.lib.orc_cluster_extents(_ptr(filt), len(filt), _ptr(out))
        return out[:len(filt)].copy()

    def angle_sort(self, cpts, idx):
        cpts = cpts.copy()
        idx = idx.copy()
        self.lib.orc_angle_sort(_ptr(cpts), _ptr(idx), len(cpts))
        return cpts, idx

    def linefit(self, cpts, gray):
        h, w = gray.shape
        out = np.empty(max(len(cpts), 1), R.LINE_FIT_POINT)
        self._f("linefit")(_ptr(cpts), len(cpts), _ptr(np.ascontiguousarray(gray)), w, h, _ptr(out))
        return out[:len(cpts)].copy()

    def linefit_prefix(self, idx, lf):
        out = np.empty(max(len(lf), 1), R.LINE_FIT_POINT)
        self.lib.orc_linefit_prefix(_ptr(idx), _ptr(lf), len(lf), _ptr(out))
        return out[:len(lf)].copy()

    def fit_lines(self, lfp, idx, ext):
        out = np.empty(max(len(lfp), 1), R.CORNER)
        self._f("fit_lines")(_ptr(lfp), _ptr(idx), _ptr(ext), len(lfp), _ptr(out))
        return out[:len(lfp)].copy()

    def corner_segments(self, corners):
        n = len(corners)
        cc = np.empty(max(n, 1), R.CORNER)
        seg = np.empty(max(n, 1), R.PEAK_EXTENTS)
        nc = C.c_int64(0)
        e = self.lib.orc_corner_segments(_ptr(corners), n, _ptr(cc), _ptr(seg), C.byref(nc))
        return cc[:nc.value].copy(), seg[:e].copy()

    def do_indexing(self, seg, cc, lfp, cext):
        out = np.empty(max(len(seg), 1), R.FITTED_QUAD)
        oob = self._f("do_indexing")(_ptr(seg), len(seg), _ptr(cc), _ptr(lfp), len(lfp),
                                       _ptr(cext), _ptr(out))
        return out[:len(seg)].copy(), oob

    def process_quads(self, fitted):
        out = np.empty(max(len(fitted), 1), R.QUAD_CORNERS)
        m = self.lib.orc_process_quads(_ptr(fitted), len(fitted), _ptr(out))
        return out[:m].copy()

    # ---- whole pipeline ----------------------------------------------------
    def pipeline(self, img, f=1):
        """Runs every stage and returns all intermediates (stage-by-stage parity)."""
        out = {}
        gray = self.decimate(img, f) if f > 1 else np.ascontiguousarray(img)
        h, w = gray.shape
        out["gray"] = gray
        out["extrema"], out["filtered_extrema"], out["thresholded"] = self.threshold(gray)
        out["labels"], out["sizes"] = self.segmentation(out["thresholded"])
        out["sparse_points"], out["sparse_keys"] = self.find_boundaries(out["labels"], out["sizes"])
        out["points"], out["keys"] = self.compact_sort(out["sparse_points"], out["sparse_keys"])
        out["bounds_all"] = self.cluster_bounds(out["points"], out["keys"])
        out["bounds"] = self.filter_bounds(out["bounds_all"], w, h)
        cpts, cidx = self.cluster_points(out["points"], out["bounds"])
        out["extents"] = self.cluster_extents(out["bounds"])
        out["cluster_points"], out["cluster_index"] = self.angle_sort(cpts, cidx)
        lf = self.linefit(out["cluster_points"], gray)
        out["line_fit"] = self.linefit_prefix(out["cluster_index"], lf)
        out["sparse_corners"] = self.fit_lines(out["line_fit"], out["cluster_index"], out["extents"])
        out["corners"], out["peak_extents"] = self.corner_segments(out["sparse_corners"])
        out["fitted"], out["oob_reads"] = self.do_indexing(out["peak_extents"], out["corners"],
                                                           out["line_fit"], out["extents"])
        out["quads"] = self.process_quads(out["fitted"])
        return out

    def detect(self, img, f=1, nthreads=0, max_out=1 << 16):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w = img.shape
        out = np.empty(max_out, R.QUAD_CORNERS)
        st = Stats()
        n = self._f("detect")(_ptr(img), w, h, f, nthreads, _ptr(out), max_out, C.byref(st))
        if n < 0:
            raise ValueError(f"{w}x{h} at decimation {f}: the decimated frame must be at least "
                             "8x8 and a multiple of 4 in each dimension")
        return out[:min(n, max_out)].copy(), st.as_dict()

    def max_threads(self):
        return self.lib.orc_max_threads()


def grey_from_pixels(img):
    """Grey image of interleaved 8-bit pixels (H x W x C, C in {1, 3, 4}) as
    the reference's loader makes it: stbi_load(..., STBI_grey)
    (src/main.cpp:265) converts 3- and 4-channel pixels with stbi__compute_y,
    (77 r + 150 g + 29 b) >> 8, dropping alpha (include/stb_image.h:1834-1836,
    :1897-1916)."""
    img = np.asarray(img, np.uint8)
    if img.ndim == 2 or img.shape[2] == 1:
        return img.reshape(img.shape[0], img.shape[1]).copy()
    r, g, b = (img[..., k].astype(np.uint32) for k in range(3))
    return ((r * 77 + g * 150 + b * 29) >> 8).astype(np.uint8)


class RefStb:
    """The reference loader's colour-to-grey conversion (oracle/_ref build)."""

    def __init__(self, path=None):
        path = path or os.path.join(_HERE, "_ref", "libref_stb.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.lib.ref_stbi_grey.restype = C.c_int
        self.lib.ref_stbi_grey.argtypes = [_P, C.c_int, C.c_int, _P]

    def grey(self, pixels, channels):
        pixels = np.ascontiguousarray(pixels, np.uint8).reshape(-1)
        n = pixels.size // channels
        out = np.empty(n, np.uint8)
        if self.lib.ref_stbi_grey(_ptr(pixels), n, channels, _ptr(out)) != 0:
            raise RuntimeError("ref_stbi_grey failed")
        return out


class RefHeaders:
    """The reference's own SYCL-free header functions (oracle/_ref build)."""

    def __init__(self, path=None):
        path = path or os.path.join(_HERE, "_ref", "libref_headers.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        sig = {
            "ref_calc_theta": (C.c_float, [C.c_uint16, C.c_uint16, C.c_float, C.c_float]),
            "ref_point_fields": (None, [C.c_uint16, C.c_uint16, _P]),
            "ref_bounds_from_point": (None, [C.c_uint16, C.c_uint16, C.c_uint32, _P]),
            "ref_reduce_bounds": (None, [_P, _P, _P]),
            "ref_bounds_cx_cy_dot": (None, [_P, _P]),
            "ref_valid_blob": (C.c_int, [_P, _I, _I]),
            "ref_initial_linefit": (None, [C.c_uint16, C.c_uint16, C.c_uint32, C.c_uint32, _P, _P]),
            "ref_n_choose_k": (C.c_int, [C.c_int, C.c_int]),
            "ref_combinadics_decode": (None, [C.c_int, _P]),
            "ref_sizeof": (C.c_int, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


def ref_headers_available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libref_headers.so"))


def line_fp64(m):
    """A FittedQuad moment (LINE_FIT_POINT record) fitted in exact integer /
    float64 arithmetic: the centroid (Mx / 2W, My / 2W) and the unit normal
    = the eigenvector of the smallest eigenvalue of the weighted covariance
    [[Mxx W - Mx^2, Mxy W - Mx My], [.., Myy W - My^2]] (numpy eigh), the
    line fit_line approximates in float32 (src/fit_quad.cpp:37-105). The
    covariance entries are formed with Python integers, so nothing wraps."""
    W, Mx, My = int(m["W"]), int(m["Mx"]), int(m["My"])
    cxx = int(m["Mxx"]) * W - Mx * Mx
    cxy = int(m["Mxy"]) * W - Mx * My
    cyy = int(m["Myy"]) * W - My * My
    cov = np.array([[cxx, cxy], [cxy, cyy]], dtype=np.float64)
    _, vec = np.linalg.eigh(cov)
    n = vec[:, 0]
    return np.array([Mx / (2.0 * W), My / (2.0 * W)]), n / np.hypot(n[0], n[1])


def quad_corners_fp64(fitted_quad):
    """float64 twin of process_fitted_quads' corners (src/main.cpp:1052-1096)
    for one FITTED_QUAD record: corner i is the intersection of lines i and
    i + 1, each line solved as n . p = n . c. Returns a 4 x 2 array, or None
    when two consecutive lines are parallel to float64 precision."""
    lines = [line_fp64(fitted_quad["moments"][i]) for i in range(4)]
    out = np.empty((4, 2))
    for i in range(4):
        (c0, n0), (c1, n1) = lines[i], lines[(i + 1) % 4]
        A = np.array([n0, n1])
        if abs(np.linalg.det(A)) < 1e-12:
            return None
        out[i] = np.linalg.solve(A, np.array([n0 @ c0, n1 @ c1]))
    return out


def corner_error_fp64(fitted, quads):
    """Largest |float32 corner - float64 twin| (pixels) over `quads`
    (QUAD_CORNERS records, blob_index = their fitted quad) and the number of
    quads compared."""
    worst = 0.0
    for q in quads:
        ref = quad_corners_fp64(fitted[int(q["blob_index"])])
        if ref is None:
            continue
        worst = max(worst, float(np.abs(q["corners"].astype(np.float64) - ref).max()))
    return worst, len(quads)
