"""The stage drop-ins and glue entry points chained as the reference's main()
chains its stages (src/main.cpp:322-1023): each C-ABI call in turn, every
intermediate array copied back. Used by the parity tests of the drop-ins and
by the dump tool (pyatq.dump)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import P, check, dev_empty, dev_from_np, lib, stream, sync, to_np
from . import decimate_threshold, find_boundaries, image_segmentation
from . import records as R


def _count(buf, field):
    sync()
    return int(buf.download(R.COUNTS, 1)[0][field])


def run(gray_dev, w_in, h_in, f=1):
    """Runs threshold -> quads through the drop-ins on a grey frame in device
    memory; returns a dict of numpy arrays named as the oracle's pipeline()
    names them, plus the sparse planes and the thresholded image."""
    L = lib()
    (w, h), grey, ext, fil, thr = decimate_threshold(gray_dev, w_in, h_in, f)
    grey = grey if f > 1 else gray_dev
    labels, sizes = image_segmentation(thr, w, h)
    pts, keys = find_boundaries(labels, sizes, w, h)
    n = 4 * w * h
    ws_bytes = L.atq_glue_workspace_size(w, h)
    ws = dev_empty(ws_bytes)
    out = {"dims": (w, h), "thresholded": to_np(thr, np.uint8, w * h).reshape(h, w),
           "labels": to_np(labels, np.uint32, w * h).reshape(h, w),
           "sizes": to_np(sizes, np.uint32, w * h),
           "sparse_points": to_np(pts, np.uint32, n), "sparse_keys": to_np(keys, np.uint64, n)}
    g_pts, g_keys, cnt = dev_empty(4 * n), dev_empty(8 * n), dev_empty(4)
    check(L.atq_group_boundaries(stream(), P(pts), P(keys), w, h, P(g_pts), P(g_keys), P(cnt),
                                 P(ws), ws_bytes))
    sync()
    nbp = int(cnt.download(np.uint32, 1)[0])
    out["points"] = to_np(g_pts, np.uint32, nbp)
    out["keys"] = to_np(g_keys, np.uint64, nbp)
    cap_cl = n // 24 + 1
    counts = dev_from_np(np.zeros(1, R.COUNTS))
    bounds = dev_empty(32 * cap_cl)
    check(L.atq_cluster_bounds_filter(stream(), P(g_pts), P(g_keys), P(cnt), w, h, P(bounds),
                                      P(counts), P(ws), ws_bytes))
    ncl = _count(counts, "n_clusters")
    out["n_pairs"] = _count(counts, "n_pairs")
    out["bounds"] = to_np(bounds, R.CLUSTER_BOUNDS, ncl)
    cpts, cidx, cext = dev_empty(8 * n), dev_empty(4 * n), dev_empty(8 * cap_cl)
    check(L.atq_cluster_points_sort(stream(), P(g_pts), P(cnt), P(bounds), P(counts), P(cpts),
                                    P(cidx), P(cext), P(ws), ws_bytes))
    nfp = _count(counts, "n_cluster_points")
    out["cluster_points"] = to_np(cpts, R.CLUSTER_POINT, nfp)
    out["cluster_index"] = to_np(cidx, np.uint32, nfp)
    out["extents"] = to_np(cext, R.CLUSTER_EXTENTS, ncl)
    lf = dev_empty(64 * max(nfp, 1))
    check(L.atq_linefit_prefix(stream(), P(cpts), P(cext), P(counts), P(grey), w, h, P(lf)))
    out["line_fit"] = to_np(lf, R.LINE_FIT_POINT, nfp)
    corners = dev_empty(12 * max(nfp, 1))
    check(L.atq_fit_lines(stream(), P(lf), P(cidx), P(cext), nfp, P(corners)))
    out["sparse_corners"] = to_np(corners, R.CORNER, nfp)
    comp, segs = dev_empty(12 * max(nfp, 1)), dev_empty(8 * cap_cl)
    check(L.atq_corner_segments(stream(), P(corners), P(counts), P(comp), P(segs), P(ws), ws_bytes))
    nc, ns = _count(counts, "n_corners"), _count(counts, "n_segments")
    out["corners"] = to_np(comp, R.CORNER, nc)
    out["peak_extents"] = to_np(segs, R.PEAK_EXTENTS, ns)
    fitted = dev_empty(272 * max(ns, 1))
    check(L.atq_do_indexing(stream(), P(segs), ns, P(comp), P(lf), nfp, P(cext), P(fitted)))
    out["fitted"] = to_np(fitted, R.FITTED_QUAD, ns)
    quads, nq = dev_empty(40 * max(ns, 1)), dev_empty(4)
    check(L.atq_quads_from_fits(stream(), P(fitted), ns, P(quads), P(nq), P(ws), ws_bytes))
    sync()
    out["quads"] = to_np(quads, R.QUAD_CORNERS, int(nq.download(np.uint32, 1)[0]))
    return out
