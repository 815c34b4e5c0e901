import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sycl-apriltag-test_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.Oracle()


def load_png_gray(name):
    from PIL import Image
    im = Image.open(os.path.join(GOLDEN, name)).convert("L")
    return np.asarray(im, dtype=np.uint8).copy()


# The reference's three input images (no expected outputs ship with them).
REF_PNGS = ["ref_decimate_420x312.png", "ref_decimate4_960x540.png",
            "ref_decimate2_1920x1080.png"]
