"""The C-ABI drop-ins chained as the reference's main() chains its stages
(decimate/threshold, segmentation, find_boundaries, then the glue entry
points atq_group_boundaries ... atq_quads_from_fits), every array compared
with the oracle's element by element. Unlike the detector, the glue keeps the
reference's full intermediates (all sparse planes, every grouped record,
every pair's bounds before the filter)."""
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
           ("plain_1080p_f2", synth.plain_1080p(2001).image, 2)]
    for name in REF_PNGS[:2]:
        out.append((name, load_png_gray(name), 1))
    return out


FRAMES = _frames()


def _bytes(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("name,img,f", FRAMES, ids=[x[0] for x in FRAMES])
def test_glue_chain_matches_oracle(atq, oracle, name, img, f):
    from pyatq import glue
    h_in, w_in = img.shape
    g = glue.run(atq.dev_from_np(img), w_in, h_in, f)
    o = oracle.pipeline(img, f)
    np.testing.assert_array_equal(g["thresholded"], o["thresholded"])
    np.testing.assert_array_equal(g["labels"], o["labels"])
    roots = np.unique(o["labels"].reshape(-1) & 0x7FFFFFFF)
    np.testing.assert_array_equal(g["sizes"][roots], o["sizes"][roots])
    np.testing.assert_array_equal(g["sparse_points"], o["sparse_points"].reshape(-1))
    np.testing.assert_array_equal(g["sparse_keys"], o["sparse_keys"].reshape(-1))
    np.testing.assert_array_equal(g["points"], o["points"])
    np.testing.assert_array_equal(g["keys"], o["keys"])
    assert g["n_pairs"] == len(o["bounds_all"])
    for k in ("bounds", "cluster_points", "cluster_index", "extents", "line_fit",
              "sparse_corners", "corners", "peak_extents", "fitted", "quads"):
        np.testing.assert_array_equal(_bytes(g[k]), _bytes(o[k]), err_msg=k)
