"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle san` builds oracle/_san/oracle_san: the oracle's source and a
small driver (oracle/san_driver.cpp) in one executable, every error fatal. It
runs the whole detector (compat and fixed mode, decimation 1 and 2) on each
image below; a clean exit and the same quads as the in-process oracle are
required. Frames whose decimated dims are not multiples of 4 are rejected
(-1) without reading them.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import REF_PNGS, ROOT, load_png_gray

SAN = os.path.join(ROOT, "oracle", "_san", "oracle_san")


def _images():
    from pyatq import synth
    rng = np.random.default_rng(11)
    maze = synth.make_frame(9, 1104, 1104, 4, maze=True, blur=0.0, noise=8.0).image
    return [
        ("plain_small", synth.plain_small().image),
        ("plain_1080p", synth.plain_1080p(2000).image),
        ("diff_frame", synth.diff_frame(5000).image),
        ("maze_1104", maze),
        ("grid_1024x768", synth.square_grid(1024, 768).image),
        ("noise_8x8", rng.integers(0, 256, (8, 8), dtype=np.uint8)),
        ("noise_20x12", rng.integers(0, 256, (12, 20), dtype=np.uint8)),
        ("noise_104x68", rng.integers(0, 256, (68, 104), dtype=np.uint8)),
        ("flat_32x32", np.full((32, 32), 100, np.uint8)),
    ] + [(n, load_png_gray(n)) for n in REF_PNGS]


def _write_pgm(path, img):
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(img, np.uint8).tobytes())


def _fnv(q):
    h = 0xcbf29ce484222325
    for rec in q:
        b = (rec["corners"].astype(np.float32).tobytes() +
             np.uint8(rec["reversed_border"]).tobytes() + np.uint32(rec["blob_index"]).tobytes())
        for c in b:
            h = ((h ^ c) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.fixture(scope="module")
def san_binary():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return SAN


def test_oracle_is_clean_under_asan_and_ubsan(san_binary, oracle, tmp_path):
    imgs = _images()
    paths = []
    for name, img in imgs:
        p = str(tmp_path / f"{name}.pgm")
        _write_pgm(p, img)
        paths.append(p)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="1")
    r = subprocess.run([san_binary] + paths, capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stderr[-6000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-6000:]
    lines = r.stdout.strip().split("\n")
    assert len(lines) == 4 * len(imgs)
    got = {}
    for ln in lines:
        f = ln.split()
        got[(f[0], f[1], f[2])] = (int(f[4]), int(f[6]), int(f[8]), int(f[10], 16))
    # the same counts and quads as the in-process (unsanitized) oracle
    import pyoracle
    fixed = pyoracle.Oracle(fixed=True)
    for (name, img), p in zip(imgs, paths):
        for mode, fn in (("compat", oracle.detect), ("fixed", fixed.detect)):
            for fac in (1, 2):
                n, ncl, npts, h = got[(p, mode, f"f{fac}")]
                try:
                    q, st = fn(img, fac)
                except ValueError:  # dims the detector rejects
                    assert n == -1, (name, mode, fac)
                    continue
                assert n == len(q) and ncl == st["n_clusters"] and npts == st["n_cluster_points"], \
                    (name, mode, fac)
                assert h == _fnv(q), (name, mode, fac)
