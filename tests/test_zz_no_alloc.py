"""atq_image_segmentation_static (no allocation, library-held scratch)
against atq_image_segmentation, through the C ABI, bit-exact: labels
everywhere, sizes at every label root and at index 0. Two calls in a row,
the second on a smaller frame (the scratch keeps the first call's words),
then the first frame again, and the capacity limit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def atq():
    import pyatq
    pyatq.init()
    return pyatq


def _thresholded(atq, img):
    h, w = img.shape
    _, _, thr = atq.threshold_image(atq.dev_from_np(img), w, h)
    return thr, w, h


def _static(atq, thr, w, h):
    labels = atq.dev_empty(4 * w * h)
    sizes = atq.dev_empty(4 * w * h)
    rc = atq.lib().atq_image_segmentation_static(atq.stream(), atq.P(thr), atq.P(labels),
                                                  atq.P(sizes), w * h, w, h)
    atq.check(rc)
    atq.sync()
    return atq.to_np(labels, np.uint32, w * h), atq.to_np(sizes, np.uint32, w * h)


def _with_scratch(atq, thr, w, h):
    labels, sizes = atq.image_segmentation(thr, w, h)
    atq.sync()
    return atq.to_np(labels, np.uint32, w * h), atq.to_np(sizes, np.uint32, w * h)


def _same(a, b):
    (la, sa), (lb, sb) = a, b
    np.testing.assert_array_equal(la, lb)
    roots = np.unique(np.concatenate([[0], la & 0x7FFFFFFF]))
    np.testing.assert_array_equal(sa[roots], sb[roots])


def _frames():
    from pyatq import synth
    big = synth.plain_1080p(2000).image                       # 1920x1080
    small = synth.plain_small().image[::2, ::2].copy()        # 320x240
    return big, small


def test_static_segmentation_matches_scratch_entry_point(atq, oracle):
    big, small = _frames()
    tb = _thresholded(atq, big)
    ts = _thresholded(atq, small)
    want_big = _with_scratch(atq, *tb)
    want_small = _with_scratch(atq, *ts)
    print("static segmentation: 1920x1080, then 320x240, then 1920x1080", flush=True)
    _same(_static(atq, *tb), want_big)
    _same(_static(atq, *ts), want_small)
    _same(_static(atq, *tb), want_big)
    # the small frame's labels are the oracle's too
    lab, _ = oracle.segmentation(oracle.threshold(small)[2])
    np.testing.assert_array_equal(_static(atq, *ts)[0].reshape(lab.shape), lab)


def test_static_segmentation_rejects_frames_above_its_capacity(atq):
    cap = 4194304  # ATQ_SEGMENTATION_STATIC_MAX_PIXELS
    w, h = 4096, 1028  # 4,210,688 pixels > 4,194,304
    assert w * h > cap
    dummy = atq.dev_empty(16)
    rc = atq.lib().atq_image_segmentation_static(atq.stream(), atq.P(dummy), atq.P(dummy),
                                                  atq.P(dummy), 0, w, h)
    assert rc == 2  # ATQ_ERR_CAPACITY; nothing is launched
