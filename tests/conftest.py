import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def tpt():
    return importlib.import_module("tiny-path-tracer_amd")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle_lib import REF_SO, Reference

    if not os.path.exists(REF_SO):
        pytest.skip("reference build oracle/_ref/libtpt_ref.so not present (needs the reference tree at build time)")
    return Reference()


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    return np.load(os.path.join(ROOT, "tests", "golden", "ref_hits.npz"))


@pytest.fixture(scope="session")
def device(tpt):
    import torch

    if tpt.lib().tpt_device_count() < 1 or not torch.cuda.is_available():
        pytest.fail("GPU test without a visible gfx950 device")
    return torch.device("cuda:0")


@pytest.fixture
def earth_cwd(tmp_path, monkeypatch):
    """A working directory holding an earthmap.jpg (the scenes that use the
    earth texture load it from there, as the reference does): a seeded
    1024x512 baseline JPEG written with PIL."""
    import numpy as np

    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(2024)
    y, x = np.mgrid[0:512, 0:1024] / 512.0
    img = np.stack([127 + 100 * np.sin(3 * x + c) * np.cos(5 * y - c) for c in (0.0, 1.0, 2.0)], -1)
    img = np.clip(img + rng.normal(0, 8, img.shape), 0, 255).astype(np.uint8)
    Image.fromarray(img).save(tmp_path / "earthmap.jpg", "JPEG", quality=90, subsampling=0)
    monkeypatch.chdir(tmp_path)
    return tmp_path
