"""The detector's fixed mode on the GPU against the oracle's fixed mode:
every intermediate and the quads, bit-exact, as for the compat mode
(tests/test_gpu_pipeline.py); plus the labels and sizes of the fixed CCL.
"""
import numpy as np
import pytest

from test_gpu_pipeline import FRAMES, _bytes, check_against_oracle, run_detector

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def atq():
    import pyatq
    pyatq.init()
    return pyatq


@pytest.fixture(scope="module")
def fixed_oracle():
    import pyoracle
    return pyoracle.Oracle(fixed=True)


def _fixed_frames():
    from pyatq import synth
    return FRAMES + [("diff_frame", synth.diff_frame().image, 1)]


FIXED_FRAMES = _fixed_frames()


@pytest.mark.parametrize("name,img,f", FIXED_FRAMES, ids=[x[0] for x in FIXED_FRAMES])
def test_fixed_detector_matches_fixed_oracle(atq, fixed_oracle, name, img, f):
    o = fixed_oracle.pipeline(img, f)
    g = run_detector(atq, img, f, fixed=True)
    check_against_oracle(g, o)
    q, _ = fixed_oracle.detect(img, f)
    np.testing.assert_array_equal(_bytes(g["quads"]), _bytes(q))


def test_fixed_labels_and_sizes(atq, fixed_oracle):
    """Labels bit-exact; sizes exact below the detector's saturation point
    (min_component_size) and on the same side of it above; sizes[0] = 0."""
    from pyatq import synth
    img = synth.plain_1080p(2006).image
    h, w = img.shape
    labels_o, sizes_o = fixed_oracle.segmentation(fixed_oracle.threshold(img)[2])
    det = atq.Detector(w, h, fixed=True)
    det.detect(atq.dev_from_np(img), w, h)
    atq.sync()
    labels = det.read("LABELS", np.uint32, w * h)
    sizes = det.read("SIZES", np.uint32, w * h + 1)
    det.close()
    np.testing.assert_array_equal(labels, labels_o.reshape(-1))
    assert sizes[0] == 0
    roots = np.unique(labels_o & 0x7FFFFFFF)
    roots = roots[roots != 0]
    so, sg = sizes_o[roots], sizes[roots]
    np.testing.assert_array_equal(sg >= 25, so >= 25)
    np.testing.assert_array_equal(sg[so < 25], so[so < 25])


def test_fixed_detector_stress_12mp(atq, fixed_oracle):
    from pyatq import synth
    img = synth.stress_12mp().image
    o = fixed_oracle.pipeline(img, 1)
    g = run_detector(atq, img, 1, fixed=True)
    check_against_oracle(g, o)


def test_fixed_detector_pair_table_overflow(atq, fixed_oracle):
    yy, xx = np.mgrid[0:480, 0:640]
    img = np.where(((yy % 7) < 5) & ((xx % 7) < 5), 220, 30).astype(np.uint8)
    o = fixed_oracle.pipeline(img, 1)
    g = run_detector(atq, img, 1, pair_table_log2=8, fixed=True)
    assert int(g["counts"]["pair_filter_off"]) == 1
    check_against_oracle(g, o)


def test_compat_and_fixed_detectors_side_by_side(atq, oracle, fixed_oracle):
    """Two detectors of different modes on one frame keep their own results."""
    from pyatq import synth
    img = synth.plain_1080p(2007).image
    h, w = img.shape
    dev = atq.dev_from_np(img)
    dc, df = atq.Detector(w, h), atq.Detector(w, h, fixed=True)
    dc.detect(dev, w, h)
    df.detect(dev, w, h)
    atq.sync()
    qc, _ = oracle.detect(img, 1)
    qf, _ = fixed_oracle.detect(img, 1)
    np.testing.assert_array_equal(_bytes(dc.quads()), _bytes(qc))
    np.testing.assert_array_equal(_bytes(df.quads()), _bytes(qf))
    assert len(qf) != len(qc)
    dc.close()
    df.close()
