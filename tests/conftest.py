"""Shared test setup: import paths, the ``gpu`` marker and golden fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for sub in ("cuda-path-tracer_amd", "oracle"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libptrace.so")


def golden(name):
    """(scene, camera, extra arrays) of a fixture; scene is None for data-only files."""
    import ptrace
    path = os.path.join(GOLDEN, name)
    with np.load(path, allow_pickle=False) as z:
        if "triangles" not in z.files:
            return None, None, {k: z[k] for k in z.files}
    return ptrace.load_scene(path)


@pytest.fixture(scope="session")
def golden_loader():
    return golden


def result_bytes(a):
    return np.ascontiguousarray(a).view(np.uint8).reshape(len(a), -1)
