"""The dump tool (pyatq.dump): the reference's CSV schemas and ostream
formatting (src/main.cpp:51-237) on the CPU; on the GPU, every file of one
frame against the oracle's arrays written by the same writers."""
import os

import numpy as np
import pytest


def _read(path):
    with open(path) as f:
        return f.read().splitlines()


def test_writers_follow_the_reference_schemas(tmp_path):
    from pyatq import dump
    from pyatq import records as R
    c = np.zeros(3, R.CORNER)
    c["error"] = [0.1234567, 65536.5, 1e-5]
    c["line_fit_point_index"] = [7, 8, 9]
    c["cluster_index"] = [1, 2, 3]
    dump.write_corners(tmp_path / "c.csv", c)
    assert _read(tmp_path / "c.csv") == ["error,line_fit_point_index,cluster_index",
                                          "0.123457,7,1", "65536.5,8,2", "1e-05,9,3"]
    # packed_x = x | plane << 14, packed_y = y | black-to-white << 15
    packed = np.array([5 | (3 << 14) | ((9 | (1 << 15)) << 16), 0], np.uint32)
    dump.write_boundary_points(tmp_path / "b.csv", packed, np.array([12345, 2**64 - 1], np.uint64))
    assert _read(tmp_path / "b.csv") == ["x_value,y_value,blob_label,is_black_to_white,dx,dy",
                                          "5,9,12345,1,1,0", "0,0,18446744073709551615,0,1,0"]
    e = np.zeros(2, R.CLUSTER_EXTENTS)
    e["start"], e["count"] = [0, 40], [40, 24]
    dump.write_extents(tmp_path / "e.csv", e)
    assert _read(tmp_path / "e.csv") == ["start,count", "0,40", "40,24"]
    lf = np.zeros(1, R.LINE_FIT_POINT)
    lf["Mx"], lf["Mxx"], lf["W"] = -3, 2**40, 7
    dump.write_line_fit(tmp_path / "l.csv", lf)
    assert _read(tmp_path / "l.csv") == ["Mx,My,Mxx,Myy,Mxy,W", f"-3,0,{2**40},0,0,7"]
    dump.write_plain(tmp_path / "p.csv", np.array([[1, 2], [3, 4]], np.uint32))
    assert _read(tmp_path / "p.csv") == ["data", "1", "2", "3", "4"]


@pytest.mark.gpu
def test_dump_frame_matches_oracle(tmp_path, oracle):
    from pyatq import dump, synth
    img = synth.plain_small().image
    g_dir, o_dir = tmp_path / "gpu", tmp_path / "oracle"
    dump.dump_frame(img, 2, str(g_dir))
    o = oracle.pipeline(img, 2)
    os.makedirs(o_dir)
    P = lambda n: str(o_dir / n)  # noqa: E731
    dump.write_plain(P("outneg2.csv"), o["labels"])
    dump.write_boundary_points(P("out0.csv"), o["sparse_points"].reshape(-1), o["sparse_keys"].reshape(-1))
    keep = o["sparse_points"].reshape(-1) != 0
    dump.write_boundary_points(P("out1.csv"), o["sparse_points"].reshape(-1)[keep],
                               o["sparse_keys"].reshape(-1)[keep])
    dump.write_boundary_points(P("out2.csv"), o["points"], o["keys"])
    dump.write_cluster_bounds(P("out2bounds.csv"), o["bounds_all"])
    cpts, cidx = oracle.cluster_points(o["points"], o["bounds"])
    dump.write_cluster_points(P("out3.csv"), cpts)
    dump.write_plain(P("out4.csv"), cidx)
    dump.write_extents(P("out4extents.csv"), o["extents"])
    dump.write_cluster_points(P("out5.csv"), o["cluster_points"])
    dump.write_line_fit(P("out6a.csv"), oracle.linefit(o["cluster_points"], o["gray"]))
    dump.write_line_fit(P("out6b.csv"), o["line_fit"])
    sc = o["sparse_corners"]
    dump.write_corners(P("out7.csv"), sc[sc["error"] != 0])
    dump.write_corners(P("out8.csv"), o["corners"])
    dump.write_extents(P("out7extents.csv"), o["peak_extents"])
    for name in sorted(os.listdir(o_dir)):
        assert _read(g_dir / name) == _read(o_dir / name), name
    assert (g_dir / "outneg3.csv").exists()
