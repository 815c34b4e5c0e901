"""Debug dumps in the reference's CSV schemas (src/main.cpp:51-237), written
from this library's GPU results, for cross-checking against a SYCL build of
the reference. One frame goes through the C-ABI drop-ins (pyatq.glue) and
every array the reference dumps in debug mode is written under the same file
name (src/main.cpp:370-372, 429, 492, 534, 599, 694-696, 731, 764, 789, 817,
851, 878, 962):

  outneg2 labels, outneg3 sizes, out0 sparse boundary planes, out1 compacted
  records, out2 grouped records, out2bounds every pair's ClusterBounds,
  out3 / out4 valid clusters' points (before the angle sort) and cluster
  indices, out4extents, out5 angle-sorted points, out6a line-fit points,
  out6b their prefix sums, out7 compacted corners, out8 sorted corners,
  out7extents peak extents.

Not written: outneg1 (the reference's union-find scratch array, whose
contents belong to its own kernel organisation). Values are formatted as a
C++ ostream formats them by default (integers in full, floats with 6
significant digits, bools as 0/1). In outneg3 only the entries at label
roots (and index 0) are defined, in the reference and here.

Usage: python -m pyatq.dump FRAME.png|FRAME.npy [--decimate F] [--out DIR]
(colour PNGs are converted to grey as the reference's loader does).
"""
from __future__ import annotations

import argparse
import os

import numpy as np

COORD_VALUE_MASK = (1 << 14) - 1
COLOR_DIRECTION_MASK = 1 << 15
HALF_PIXEL_MASK = (1 << 15) & (1 << 14)  # zero, as in include/boundary_point.h:16
# dx, dy of the four half-pixel codes (include/boundary_point.h:63-95)
_DX = np.array([1, -1, 0, 1], np.int64)
_DY = np.array([0, 1, 1, 1], np.int64)


def _g(x):
    return "%g" % float(x)


def _point_fields(packed):
    """x_value, y_value, is_black_to_white, dx, dy of packed u32 records
    (packed_x in the low half, packed_y in the high half)."""
    px = (packed & 0xFFFF).astype(np.int64)
    py = (packed >> 16).astype(np.int64)
    hp = (px & HALF_PIXEL_MASK) >> 14
    return (px & COORD_VALUE_MASK, py & COORD_VALUE_MASK, (py & COLOR_DIRECTION_MASK) != 0,
            _DX[hp], _DY[hp])


def _write(path, header, rows):
    with open(path, "w") as f:
        f.write(header + "\n")
        for r in rows:
            f.write(",".join(r) + "\n")


def write_plain(path, data):
    _write(path, "data", ([str(int(v))] for v in np.asarray(data).reshape(-1)))


def write_boundary_points(path, packed, labels):
    x, y, b2w, dx, dy = _point_fields(np.asarray(packed, np.uint32))
    labels = np.asarray(labels, np.uint64)
    _write(path, "x_value,y_value,blob_label,is_black_to_white,dx,dy",
           ([str(x[i]), str(y[i]), str(int(labels[i])), str(int(b2w[i])), str(dx[i]), str(dy[i])]
            for i in range(len(labels))))


def write_cluster_bounds(path, b):
    f = ("x_min", "x_max", "y_min", "y_max", "gx_sum", "gy_sum", "pxgx_pygy_sum", "start", "count")
    _write(path, ",".join(f), ([str(int(r[k])) for k in f] for r in b))


def write_cluster_points(path, pts):
    packed = pts["packed_x"].astype(np.uint32) | (pts["packed_y"].astype(np.uint32) << 16)
    x, y, b2w, dx, dy = _point_fields(packed)
    _write(path, "x_value,y_value,is_black_to_white,dx,dy,slope",
           ([str(x[i]), str(y[i]), str(int(b2w[i])), str(dx[i]), str(dy[i]), _g(pts["slope"][i])]
            for i in range(len(pts))))


def write_line_fit(path, lf):
    f = ("Mx", "My", "Mxx", "Myy", "Mxy", "W")
    _write(path, ",".join(f), ([str(int(r[k])) for k in f] for r in lf))


def write_corners(path, c):
    _write(path, "error,line_fit_point_index,cluster_index",
           ([_g(r["error"]), str(int(r["line_fit_point_index"])), str(int(r["cluster_index"]))]
            for r in c))


def write_extents(path, e):
    _write(path, "start,count", ([str(int(r["start"])), str(int(r["count"]))] for r in e))


def _bounds_all(points, keys):
    """Every pair's ClusterBounds (include/cluster_bounds.h:14-83) from the
    grouped records: the reference's reduce_by_segment output."""
    from . import records as R
    n = len(keys)
    heads = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]]) if n else np.zeros(0, np.int64)
    x, y, b2w, dx, dy = _point_fields(points)
    xx, yy = 2 * x + dx, 2 * y + dy
    gx, gy = np.where(b2w, dx, -dx), np.where(b2w, dy, -dy)
    out = np.zeros(len(heads), R.CLUSTER_BOUNDS)
    if not n:
        return out
    out["x_min"] = np.minimum.reduceat(xx, heads)
    out["x_max"] = np.maximum.reduceat(xx, heads)
    out["y_min"] = np.minimum.reduceat(yy, heads)
    out["y_max"] = np.maximum.reduceat(yy, heads)
    out["gx_sum"] = np.add.reduceat(gx, heads).astype(np.int32)
    out["gy_sum"] = np.add.reduceat(gy, heads).astype(np.int32)
    out["pxgx_pygy_sum"] = np.add.reduceat(xx * gx + yy * gy, heads)
    out["start"] = heads
    out["count"] = np.diff(np.r_[heads, n])
    return out


def dump_frame(gray, f, out_dir):
    """Runs one grey frame (H x W u8, host memory) through the drop-ins on the
    GPU and writes the reference's debug CSVs into out_dir."""
    from . import dev_from_np, glue, init
    init()
    os.makedirs(out_dir, exist_ok=True)
    p = lambda name: os.path.join(out_dir, name)  # noqa: E731
    h_in, w_in = gray.shape
    g = glue.run(dev_from_np(np.ascontiguousarray(gray, np.uint8)), w_in, h_in, f)
    write_plain(p("outneg2.csv"), g["labels"])
    write_plain(p("outneg3.csv"), g["sizes"])
    write_boundary_points(p("out0.csv"), g["sparse_points"], g["sparse_keys"])
    keep = g["sparse_points"] != 0
    write_boundary_points(p("out1.csv"), g["sparse_points"][keep], g["sparse_keys"][keep])
    write_boundary_points(p("out2.csv"), g["points"], g["keys"])
    write_cluster_bounds(p("out2bounds.csv"), _bounds_all(g["points"], g["keys"]))
    # the valid clusters' records in grouped order, each with the slope the
    # angle sort carried (a record's packed coordinates are unique in its pair)
    cp, ci, ext = g["cluster_points"], g["cluster_index"], g["extents"]
    pre = np.empty_like(cp)
    for c, b in enumerate(g["bounds"]):
        s, n = int(ext[c]["start"]), int(ext[c]["count"])
        mine = cp[s:s + n]
        key_sorted = (mine["packed_x"].astype(np.uint32) | (mine["packed_y"].astype(np.uint32) << 16))
        grouped = g["points"][int(b["start"]):int(b["start"]) + int(b["count"])]
        order = np.argsort(key_sorted, kind="stable")
        at = order[np.searchsorted(key_sorted[order], grouped)]
        pre[s:s + n] = mine[at]
    write_cluster_points(p("out3.csv"), pre)
    write_plain(p("out4.csv"), ci)
    write_extents(p("out4extents.csv"), ext)
    write_cluster_points(p("out5.csv"), cp)
    lf = g["line_fit"]
    raw = lf.copy()
    for c in range(len(ext)):
        s, n = int(ext[c]["start"]), int(ext[c]["count"])
        for k in ("Mx", "My", "Mxx", "Myy", "Mxy", "W"):
            raw[k][s + 1:s + n] = np.diff(lf[k][s:s + n])
    write_line_fit(p("out6a.csv"), raw)
    write_line_fit(p("out6b.csv"), lf)
    sc = g["sparse_corners"]
    write_corners(p("out7.csv"), sc[sc["error"] != 0])
    write_corners(p("out8.csv"), g["corners"])
    write_extents(p("out7extents.csv"), g["peak_extents"])
    return g


def _load(path):
    if path.endswith(".npy"):
        return np.load(path), 1
    from PIL import Image
    im = Image.open(path)
    a = np.asarray(im.convert("RGBA") if im.mode in ("P", "LA", "PA") else im, np.uint8)
    return a, 1 if a.ndim == 2 else a.shape[2]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("frame")
    ap.add_argument("--decimate", type=int, default=1)
    ap.add_argument("--out", default=".")
    args = ap.parse_args(argv)
    img, channels = _load(args.frame)
    if channels != 1:  # the reference loader's grey (stbi_load, STBI_grey)
        from . import dev_from_np, grey_decimate, init, to_np
        init()
        h_in, w_in = img.shape[:2]
        (w, h), g = grey_decimate(dev_from_np(img), w_in, h_in, channels, 1)
        img = to_np(g, np.uint8, w * h).reshape(h, w)
    dump_frame(img, args.decimate, args.out)
    print(f"wrote the reference's debug CSVs to {args.out}")


if __name__ == "__main__":
    main()
