import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU test, run with MI_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("MI_SLOW"):
        return
    skip = pytest.mark.skip(reason="slow; set MI_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def ctx():
    """One device context for the GPU tests (fails, never skips, without a GPU)."""
    import medimage_amd as mi
    c = mi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    d = Path(__file__).resolve().parent / "golden"
    def load(name):
        with np.load(d / name, allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    return load
