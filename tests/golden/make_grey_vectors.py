"""Writes tests/golden/ref_grey_vectors.npz from the reference's own loader
code (oracle/_ref/libref_stb.so: stbi__convert_format of the reference
checkout's include/stb_image.h, the conversion behind
stbi_load(..., STBI_grey) at src/main.cpp:265).

Inputs: every value of each channel against fixed others, the primaries and
greys, and seeded random RGB and RGBA pixels; the file holds inputs and the
reference's outputs only.
Run: python tests/golden/make_grey_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402


def main():
    ref = pyoracle.RefStb()
    rng = np.random.default_rng(2024)
    v = np.arange(256, dtype=np.uint8)
    sweeps = []
    for k in range(3):
        for other in (0, 128, 255):
            px = np.full((256, 3), other, np.uint8)
            px[:, k] = v
            sweeps.append(px)
    grey_ramp = np.repeat(v[:, None], 3, axis=1)
    rgb = np.concatenate(sweeps + [grey_ramp, rng.integers(0, 256, (8192, 3), dtype=np.uint8)])
    rgba = rng.integers(0, 256, (8192, 4), dtype=np.uint8)
    out = {"rgb": rgb, "rgb_grey": ref.grey(rgb, 3), "rgba": rgba, "rgba_grey": ref.grey(rgba, 4)}
    path = os.path.join(ROOT, "tests", "golden", "ref_grey_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, {k: a.shape for k, a in out.items()})


if __name__ == "__main__":
    main()
