"""Tag-payload decode on the GPU (atq_decode_payloads, atq_match_codes).

The detector's quads are decoded on the GPU and checked against the CPU
restatement (oracle/decode_ref.py) on the same quads: code words bit-exact,
decision margins within 1e-4 relative, homographies within 1e-5 relative.
The rendered tags' payloads are the ground truth: every detected quad that
matches a rendered, unoccluded tag reads that tag's payload, and matching
against a codebook of the frame's payloads returns the tag with Hamming 0.
Parity with the reference's decoder (the apriltag library's
quad_decode_task, src/quad_decoding.cpp:46-75, absent here) is unpinned.
"""
import os
import sys

import numpy as np
import pytest

from conftest import load_png_gray

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import decode_ref as D  # noqa: E402
from diff_harness import match_ground_truth  # noqa: E402
from test_decode import rendered_tags  # noqa: E402

pytestmark = pytest.mark.gpu
LAYOUT = D.row_major_layout()


@pytest.fixture(scope="module")
def atq():
    import pyatq
    pyatq.init()
    return pyatq


def _frames():
    from pyatq import synth
    return [("plain_1080p", synth.plain_1080p(2000)), ("plain_1080p_b", synth.plain_1080p(2003)),
            ("diff_frame", synth.diff_frame(5000)), ("diff_frame_b", synth.diff_frame(5001))]


def _detect_and_decode(atq, img, use_count=False):
    from pyatq import records as R
    h, w = img.shape
    det = atq.Detector(w, h, decimate=1)
    dev = atq.dev_from_np(img)
    det.detect(dev, w, h)
    atq.sync()
    quads = det.quads()
    qptr, qcap = det.buffer("QUADS")
    import ctypes as C
    if use_count:
        cptr, _ = det.buffer("COUNTS")
        n_dev = cptr + R.COUNTS.fields["n_quads"][1]
        out = atq.decode_payloads(dev, w, h, C.c_void_p(qptr), qcap, LAYOUT, n_dev_ptr=n_dev)
    else:
        out = atq.decode_payloads(dev, w, h, C.c_void_p(qptr), len(quads), LAYOUT)
    atq.sync()
    codes = out.download(R.TAG_CODE, len(quads))
    det.close()
    return quads, codes, dev


@pytest.mark.parametrize("name,fr", _frames(), ids=[n for n, _ in _frames()])
def test_gpu_decode_matches_the_oracle(atq, name, fr):
    quads, codes, _ = _detect_and_decode(atq, fr.image)
    assert len(quads) > 5
    worst_m, worst_h = 0.0, 0.0
    for q, c in zip(quads, codes):
        code, margin, H = D.decode(fr.image, q["corners"].astype(np.float64), LAYOUT)
        assert int(c["code"]) == code, (name, q["corners"])
        if margin < 0:
            assert c["decision_margin"] == -1.0
            continue
        worst_m = max(worst_m, abs(float(c["decision_margin"]) - margin) / max(abs(margin), 1.0))
        worst_h = max(worst_h, float(np.abs(c["homography"].reshape(3, 3) - H).max() /
                                     np.abs(H).max()))
    print(f"{name}: {len(quads)} quads, margin rel err {worst_m:.1e}, H rel err {worst_h:.1e}")
    assert worst_m <= 1e-4 and worst_h <= 1e-5


def test_gpu_decode_reads_the_rendered_payloads(atq):
    from pyatq import synth
    frames = [synth.diff_frame(5000 + k) for k in range(3)] + [synth.plain_1080p(2000)]
    found = 0
    for fr in frames:
        visible = rendered_tags([fr])
        quads, codes, _ = _detect_and_decode(atq, fr.image)
        gt = [c for _, c, _ in visible]
        table = np.array([D.payload_code(p) for _, _, p in visible], np.uint64)
        m = match_ground_truth([q["corners"].astype(np.float64) for q in quads], gt)
        code_dev = atq.dev_from_np(codes)
        table_dev = atq.dev_from_np(table)
        res = atq.match_codes(code_dev, len(codes), table_dev, len(table), LAYOUT, 3)
        atq.sync()
        from pyatq import records as R
        got = res.download(R.TAG_MATCH, len(codes))
        for k, g in enumerate(m):
            if g is None:
                continue
            idx = [i for i, c in enumerate(gt) if np.abs(np.sort(c, 0) - np.sort(g, 0)).max() < 1e-9][0]
            assert got[k]["id"] == idx and got[k]["hamming"] == 0, (k, got[k])
            found += 1
    assert found >= 20


def test_gpu_decode_count_from_the_device(atq):
    from pyatq import synth
    fr = synth.diff_frame(5000)
    quads, codes, _ = _detect_and_decode(atq, fr.image)
    quads2, codes2, _ = _detect_and_decode(atq, fr.image, use_count=True)
    assert len(quads) == len(quads2)
    np.testing.assert_array_equal(codes["code"], codes2["code"])
    np.testing.assert_array_equal(codes["decision_margin"], codes2["decision_margin"])


def test_gpu_decode_the_reference_images(atq):
    """The reference's own 1920x1080 image: codes equal the oracle's (no
    code table for its tag family is here to say which tags they are)."""
    img = load_png_gray("ref_decimate2_1920x1080.png")
    quads, codes, _ = _detect_and_decode(atq, img)
    assert len(quads) > 10
    for q, c in zip(quads, codes):
        code, margin, _ = D.decode(img, q["corners"].astype(np.float64), LAYOUT)
        assert int(c["code"]) == code
        assert (margin < 0) == (c["decision_margin"] < 0)


def test_gpu_match_equals_the_oracle(atq):
    from pyatq import records as R
    rng = np.random.default_rng(9)
    table = rng.integers(0, 1 << 36, 64, dtype=np.int64).astype(np.uint64)
    codes = np.zeros(200, R.TAG_CODE)
    for i in range(200):
        t = int(table[rng.integers(0, 64)])
        c = D.rotate(t, LAYOUT, int(rng.integers(0, 4)))
        for b in rng.choice(36, int(rng.integers(0, 6)), replace=False):
            c ^= 1 << int(b)
        codes[i]["code"] = c
        codes[i]["decision_margin"] = -1.0 if i % 37 == 0 else 10.0
    res = atq.match_codes(atq.dev_from_np(codes), len(codes), atq.dev_from_np(table), len(table),
                          LAYOUT, 3)
    atq.sync()
    got = res.download(R.TAG_MATCH, len(codes))
    tl = [int(x) for x in table]
    for i in range(len(codes)):
        want = (-1, 0, 0) if codes[i]["decision_margin"] < 0 else \
            D.match(int(codes[i]["code"]), tl, LAYOUT, 3)
        assert (got[i]["id"], got[i]["hamming"], got[i]["rotation"]) == want, i


def test_decode_rejects_bad_layouts(atq):
    import ctypes as C
    from pyatq import synth
    fr = synth.diff_frame(5000)
    dev = atq.dev_from_np(fr.image)
    quads = atq.dev_from_np(np.zeros(1, np.dtype([("c", "u1", (40,))])))
    for bad in (dict(LAYOUT, bit_x=[], bit_y=[]), dict(LAYOUT, width_at_border=1),
                dict(LAYOUT, total_width=9), dict(LAYOUT, bit_x=[12] + LAYOUT["bit_x"][1:])):
        with pytest.raises(atq.AtqError) as e:
            atq.decode_payloads(dev, 1280, 720, quads, 1, bad)
        assert e.value.code == 4  # ATQ_ERR_BAD_ARG
    # cells not closed under the quarter turn cannot be matched by rotation
    skew = dict(LAYOUT, bit_x=LAYOUT["bit_x"][:-1], bit_y=LAYOUT["bit_y"][:-1])
    codes = atq.dev_from_np(np.zeros(1, np.dtype([("c", "u1", (48,))])))
    table = atq.dev_from_np(np.zeros(1, np.uint64))
    with pytest.raises(atq.AtqError) as e:
        atq.match_codes(codes, 1, table, 1, skew, 3)
    assert e.value.code == 4
