"""Tag-payload decode oracle (oracle/decode_ref.py), pinned on the CPU.

The reference decodes quads with the apriltag library's quad_decode_task
(src/quad_decoding.cpp:46-75), a dependency absent from the reference
checkout, so nothing of it runs here: parity with it is unpinned. The
oracle's parts are pinned by what fixes them independently:
* the homography takes the tag square's corners onto the quad's (exactly);
* the border gray model recovers a planar intensity field exactly;
* the sharpening equals scipy's zero-padded convolution with the Laplacian;
* a rendered tag, decoded at its true corners, reads the payload the
  renderer drew; the inverted image fails the polarity test, and decodes to
  the complement with a reversed-border layout;
* rotations: four quarter turns are the identity, and each equals numpy's
  rot90 of the payload grid; matching finds a codebook entry under noise.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import decode_ref as D  # noqa: E402

LAYOUT = D.row_major_layout()
MASK = (1 << 36) - 1


def test_homography_maps_the_tag_corners_onto_the_quad():
    rng = np.random.default_rng(3)
    for _ in range(50):
        c = rng.uniform(0, 500, (4, 2))
        h = D.homography(c)
        for (tx, ty), (px, py) in zip(D.tag_corners(), c):
            x, y = D.project(h, tx, ty)
            assert abs(x - px) < 1e-9 and abs(y - py) < 1e-9
    # a square quad is the affine scaling
    h = D.homography([[10, 20], [30, 20], [30, 40], [10, 40]])
    np.testing.assert_allclose(h, [[10, 0, 20], [0, 10, 30], [0, 0, 1]], atol=1e-12)


def test_plane_model_is_exact_for_a_planar_field():
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1.2, 1.2, (40, 2))
    s = [(x, y, 3.5 * x - 2.25 * y + 101.0) for x, y in pts]
    np.testing.assert_allclose(D._plane(s), [3.5, -2.25, 101.0], atol=1e-9)
    assert D._plane(s[:2]) is None                               # too few samples
    assert D._plane([(0.1 * i, 0.2 * i, 5.0) for i in range(8)]) is None  # collinear


def test_sharpening_is_the_zero_padded_laplacian():
    from scipy.ndimage import convolve
    rng = np.random.default_rng(5)
    v = rng.normal(0, 50, (10, 10)).astype(np.float32)
    k = np.array([[0, -1, 0], [-1, 4, -1], [0, -1, 0]], np.float32)
    want = convolve(v.astype(np.float64), k.astype(np.float64), mode="constant", cval=0.0)
    # the oracle's sharpen, isolated: decode() with a flat model reads v back
    sharp = np.zeros_like(v)
    for y in range(10):
        for x in range(10):
            acc = np.float32(0)
            for i in range(3):
                for j in range(3):
                    yy, xx = y + i - 1, x + j - 1
                    if 0 <= yy < 10 and 0 <= xx < 10:
                        acc = np.float32(acc + v[yy, xx] * k[i, j])
            sharp[y, x] = acc
    np.testing.assert_allclose(sharp, want, rtol=1e-5, atol=1e-3)


def _box(c, grow=1.6):
    """Bounding box of the quad grown about its centre (the white ring is
    1.25x the black square; 1.6 covers the perspective)."""
    c = np.asarray(c, np.float64)
    m = c.mean(0)
    o = m + (c - m) * grow
    return o.min(0), o.max(0)


def rendered_tags(frames):
    """(image, true corners, payload) of the tags no later-drawn tag can
    overlap (the renderer paints tags in order, so later ones cover earlier
    ones)."""
    out = []
    for fr in frames:
        n = len(fr.corners)
        for i in range(n):
            lo, hi = _box(fr.corners[i])
            hidden = False
            for j in range(i + 1, n):
                lo2, hi2 = _box(fr.corners[j])
                hidden |= bool((lo < hi2).all() and (lo2 < hi).all())
            if not hidden:
                out.append((fr.image, np.asarray(fr.corners[i], np.float64), fr.payloads[i]))
    return out


def _rendered_tags():
    from pyatq import synth
    frames = [synth.diff_frame(5000 + k) for k in range(4)] + [synth.plain_small(1000 + k)
                                                               for k in range(4)]
    return rendered_tags(frames)


def test_rendered_payloads_decode_at_their_true_corners():
    tags = _rendered_tags()
    assert len(tags) >= 30
    bad = 0
    for img, c, p in tags:
        code, margin, _ = D.decode(img, c, LAYOUT)
        assert margin > 0
        bad += code != D.payload_code(p)
    assert bad == 0


def test_polarity_and_reversed_border():
    img, c, p = _rendered_tags()[0]
    inv = (255 - img.astype(np.int32)).astype(np.uint8)
    code, margin, _ = D.decode(inv, c, LAYOUT)
    assert margin == -1.0 and code == 0
    rev = dict(LAYOUT, reversed_border=1)
    code, margin, _ = D.decode(inv, c, rev)
    assert margin > 0 and code == (~D.payload_code(p) & MASK)


def test_rotations_are_quarter_turns_of_the_payload():
    rng = np.random.default_rng(6)
    for _ in range(20):
        p = rng.integers(0, 2, (6, 6))
        code = D.payload_code(p)
        assert D.rotate(code, LAYOUT, 0) == code
        r1 = D.rotate(code, LAYOUT, 1)
        assert D.rotate(r1, LAYOUT, 3) == code
        # (x, y) -> (wb-1-y, x) on the cells is numpy's rot90 of the grid (rows = y)
        assert r1 == D.payload_code(np.rot90(p, 1)) or r1 == D.payload_code(np.rot90(p, -1))
        turns = {D.rotate(code, LAYOUT, r) for r in range(4)}
        want = {D.payload_code(np.rot90(p, k)) for k in range(4)}
        assert turns == want


def test_match_finds_codes_under_rotation_and_bit_errors():
    rng = np.random.default_rng(7)
    table = [int(x) for x in rng.integers(0, 1 << 36, 50, dtype=np.int64)]
    for i in (0, 17, 49):
        for r in range(4):
            c = D.rotate(table[i], LAYOUT, (4 - r) % 4)  # read r quarter turns off
            flips = rng.choice(36, 2, replace=False)
            for b in flips:
                c ^= 1 << int(b)
            idx, ham, rot = D.match(c, table, LAYOUT, 3)
            assert (idx, ham) == (i, 2) and D.rotate(c, LAYOUT, rot) ^ table[i] != 0
    assert D.match(table[3] ^ 0xFF, table, LAYOUT, 3)[0] == -1
