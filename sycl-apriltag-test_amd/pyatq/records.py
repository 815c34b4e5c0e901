"""numpy dtypes layout-identical to the C ABI records in include/atq.h
(which in turn match the reference structs)."""
import numpy as np

EXTREMA = np.dtype([("min", "u1"), ("max", "u1")])
BOUNDARY_POINT = np.dtype([("packed_x", "<u2"), ("packed_y", "<u2")])
CLUSTER_POINT = np.dtype([("packed_x", "<u2"), ("packed_y", "<u2"), ("slope", "<f4")])
CLUSTER_EXTENTS = np.dtype([("start", "<u4"), ("count", "<u4")])
CLUSTER_BOUNDS = np.dtype([("x_min", "<u2"), ("x_max", "<u2"), ("y_min", "<u2"),
                           ("y_max", "<u2"), ("gx_sum", "<i4"), ("gy_sum", "<i4"),
                           ("pxgx_pygy_sum", "<i8"), ("start", "<u4"), ("count", "<u4")])
LINE_FIT_POINT = np.dtype([(n, "<i8") for n in ("Mx", "My", "Mxx", "Myy", "Mxy", "W", "a", "b")])
PEAK_EXTENTS = np.dtype([("start", "<u4"), ("count", "<u4")])
CORNER = np.dtype([("line_fit_point_index", "<u4"), ("cluster_index", "<u4"), ("error", "<f4")])
FITTED_QUAD = np.dtype([("moments", LINE_FIT_POINT, (4,)), ("num_in_moments", "<u2", (4,)),
                        ("indices", "<u2", (4,))])
QUAD_CORNERS = np.dtype([("corners", "<f4", (4, 2)), ("reversed_border", "u1"),
                         ("pad_", "u1", (3,)), ("blob_index", "<u4")])
DIFF_QUAD = np.dtype([("cluster_start", "<u4"), ("cluster_count", "<u4"), ("idx", "<u4", (4,))])
COUNTS = np.dtype([(n, "<u4") for n in (
    "n_boundary", "n_pairs", "n_clusters", "n_cluster_points", "n_corners", "n_segments",
    "n_quads", "overflow", "oob_reads", "key_label_bits", "pair_filter_off",
    "n_valid_pairs", "sort_passes", "rank_shift")]
    + [("reserved", "<u4", (2,))])

TAG_LAYOUT = np.dtype([("width_at_border", "<i4"), ("total_width", "<i4"),
                       ("reversed_border", "<i4"), ("nbits", "<i4"),
                       ("decode_sharpening", "<f4"), ("bit_x", "u1", (64,)),
                       ("bit_y", "u1", (64,))])
TAG_CODE = np.dtype([("code", "<u8"), ("decision_margin", "<f4"), ("homography", "<f4", (9,))])
TAG_MATCH = np.dtype([("id", "<i4"), ("hamming", "<i4"), ("rotation", "<i4"), ("pad_", "<i4")])

SIZES = {
    "EXTREMA": 2, "BOUNDARY_POINT": 4, "CLUSTER_POINT": 8, "CLUSTER_EXTENTS": 8,
    "CLUSTER_BOUNDS": 32, "LINE_FIT_POINT": 64, "PEAK_EXTENTS": 8, "CORNER": 12,
    "FITTED_QUAD": 272, "QUAD_CORNERS": 40, "COUNTS": 64, "DIFF_QUAD": 24,
    "TAG_LAYOUT": 148, "TAG_CODE": 48, "TAG_MATCH": 16,
}
for _name, _size in SIZES.items():
    assert globals()[_name].itemsize == _size, (_name, globals()[_name].itemsize)
