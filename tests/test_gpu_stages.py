"""Stage parity on the GPU: each C-ABI drop-in against the oracle, bit-exact.

Decimate, threshold (extrema, filtered extrema, thresholded),
segmentation (labels, sizes at every label root and at index 0), boundary
sparse planes.
"""
import numpy as np
import pytest

from conftest import REF_PNGS, load_png_gray

pytestmark = pytest.mark.gpu


def _cases():
    from pyatq import synth
    rng = np.random.default_rng(7)
    cases = [
        ("plain_small_f2", synth.plain_small().image, 2),
        ("plain_1080p_f1", synth.plain_1080p(2000).image, 1),
        ("plain_1080p_f2", synth.plain_1080p(2001).image, 2),
        ("noise_64x48", rng.integers(0, 256, (48, 64), dtype=np.uint8), 1),
        ("noise_8x8", rng.integers(0, 256, (8, 8), dtype=np.uint8), 1),
        ("noise_20x12", rng.integers(0, 256, (20, 12), dtype=np.uint8), 1),
        ("blocks_96x64", np.kron(rng.integers(0, 2, (16, 24)), np.ones((4, 4))).astype(np.uint8) * 255, 1),
        ("flat_32x32", np.full((32, 32), 100, np.uint8), 1),
        ("checker_32x32", (np.indices((32, 32)).sum(0) % 2 * 255).astype(np.uint8), 1),
        ("noise_f3_106x70", rng.integers(0, 256, (70, 106), dtype=np.uint8), 3),
    ]
    for name in REF_PNGS:
        cases.append((name, load_png_gray(name), 1))
    return cases


CASES = _cases()


@pytest.fixture(scope="module")
def atq():
    import pyatq
    pyatq.init()
    return pyatq


def _up(a):
    import pyatq
    return pyatq.dev_from_np(a)


@pytest.mark.parametrize("name,img,f", CASES, ids=[c[0] for c in CASES])
def test_threshold_segmentation_boundaries(atq, oracle, name, img, f):
    h_in, w_in = img.shape
    ref = {}
    gray = oracle.decimate(img, f) if f > 1 else img
    h, w = gray.shape
    ext, fil, thr = oracle.threshold(gray)
    labels, sizes = oracle.segmentation(thr)
    pts, keys = oracle.find_boundaries(labels, sizes)

    dev_in = _up(img)
    (gw, gh), g_gray, g_ext, g_fil, g_thr = atq.decimate_threshold(dev_in, w_in, h_in, f)
    assert (gw, gh) == (w, h)
    if f > 1:
        (dw, dh), g_dec = atq.decimate(dev_in, w_in, h_in, f)
        np.testing.assert_array_equal(atq.to_np(g_dec, np.uint8, w * h).reshape(h, w), gray)
        np.testing.assert_array_equal(atq.to_np(g_gray, np.uint8, w * h).reshape(h, w), gray)
    from pyatq import records as R
    np.testing.assert_array_equal(atq.to_np(g_ext, R.EXTREMA, ext.size).reshape(ext.shape), ext)
    np.testing.assert_array_equal(atq.to_np(g_fil, R.EXTREMA, fil.size).reshape(fil.shape), fil)
    np.testing.assert_array_equal(atq.to_np(g_thr, np.uint8, w * h).reshape(h, w), thr)

    if f == 1:  # the plain drop-in, include/threshold.h:6-10
        e2, f2, t2 = atq.threshold_image(dev_in, w, h)
        np.testing.assert_array_equal(atq.to_np(t2, np.uint8, w * h).reshape(h, w), thr)
        np.testing.assert_array_equal(atq.to_np(f2, R.EXTREMA, fil.size).reshape(fil.shape), fil)

    g_labels, g_sizes = atq.image_segmentation(_up(thr), w, h)
    gl = atq.to_np(g_labels, np.uint32, w * h).reshape(h, w)
    np.testing.assert_array_equal(gl, labels)
    gs = atq.to_np(g_sizes, np.uint32, w * h)
    roots = np.unique(np.concatenate([labels.reshape(-1) & 0x7FFFFFFF, [0]]))
    np.testing.assert_array_equal(gs[roots], sizes[roots])

    g_pts, g_keys = atq.find_boundaries(g_labels, g_sizes, w, h)
    np.testing.assert_array_equal(atq.to_np(g_pts, np.uint32, 4 * w * h), pts)
    np.testing.assert_array_equal(atq.to_np(g_keys, np.uint64, 4 * w * h), keys)


def test_bad_dims_rejected(atq):
    t = _up(np.zeros((10, 10), np.uint8))
    with pytest.raises(atq.AtqError):
        atq.threshold_image(t, 10, 10)
    with pytest.raises(atq.AtqError):
        atq.image_segmentation(t, 6, 6)
    # above the 8192 limit (buffers sized for the dims, so nothing reads past them)
    for w, h in ((8196, 8), (8, 8196)):
        big = _up(np.zeros((h, w), np.uint8))
        with pytest.raises(atq.AtqError) as e:
            atq.threshold_image(big, w, h)
        assert e.value.code == 1  # ATQ_ERR_BAD_DIMS
        with pytest.raises(atq.AtqError) as e:
            atq.image_segmentation(big, w, h)
        assert e.value.code == 1
        with pytest.raises(atq.AtqError) as e:
            atq.Detector(w, h, decimate=1)
        assert e.value.code == 1
    # a frame larger than the detector was created for
    det = atq.Detector(64, 64, decimate=1)
    frame = _up(np.zeros((128, 128), np.uint8))
    with pytest.raises(atq.AtqError) as e:
        det.detect(frame, 128, 128)
    assert e.value.code == 2  # ATQ_ERR_CAPACITY
    det.close()


def test_segmentation_repeatable(atq, oracle):
    """Labels are identical across repeated runs (lock-free unions race)."""
    from pyatq import synth
    img = synth.plain_1080p(2002).image
    h, w = img.shape
    _, _, _, _, thr = atq.decimate_threshold(_up(img), w, h, 1, want_extrema=False)
    first = None
    for _ in range(5):
        g_labels, _ = atq.image_segmentation(thr, w, h)
        cur = atq.to_np(g_labels, np.uint32, w * h)
        if first is None:
            first = cur
        else:
            np.testing.assert_array_equal(cur, first)
