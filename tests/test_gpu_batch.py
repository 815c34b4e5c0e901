"""Batched form: 16 frames of uneven sizes and mixed decimation through one
graph launch, bit-identical per frame to the plain per-frame detector (its
labels and quads); the graph is replayed several times with eager work in
between."""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(n=16):
    from pyatq import synth
    return synth.batch_mixed(4000, n=n)


def test_batch_matches_plain_per_frame(oracle):
    import pyatq as atq
    atq.init()
    frames = _frames()
    sizes = [(f.image.shape[1], f.image.shape[0]) for f in frames]
    decs = [f.decimate for f in frames]
    print("batch sizes:", list(zip(sizes, decs)))
    def labels_digest(det):
        w, h = det.dims()
        return hashlib.sha1(det.read("LABELS", np.uint32, w * h).tobytes()).hexdigest()

    plain, plain_labels = [], []
    for f in frames:
        h, w = f.image.shape
        det = atq.Detector(w, h, decimate=f.decimate)
        det.detect(atq.dev_from_np(f.image), w, h)
        atq.sync()
        plain.append(det.quads().tobytes())
        plain_labels.append(labels_digest(det))
        det.close()
    b = atq.Batch(sizes, decs)
    for i, f in enumerate(frames):
        b.upload(i, f.image)
    stream = atq.Stream()
    for use_graph in (False, True, True, True):
        b.detect(stream.handle, use_graph=use_graph)
        stream.sync()
        for i in range(len(frames)):
            assert b.frame(i).quads().tobytes() == plain[i], (use_graph, i)
            assert labels_digest(b.frame(i)) == plain_labels[i], (use_graph, i)
        # eager work between replays
        det = atq.Detector(640, 480, decimate=2)
        det.detect(atq.dev_from_np(frames[0].image[:480, :640].copy()), 640, 480)
        atq.sync()
        det.close()
    # spot-check two frames against the oracle itself
    for i in (0, len(frames) - 1):
        q, _ = oracle.detect(frames[i].image, frames[i].decimate)
        assert b.frame(i).quads().tobytes() == q.tobytes()
    b.close()
