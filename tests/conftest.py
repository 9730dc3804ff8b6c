"""Shared test set-up: import paths, the `gpu` marker, common fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ww3-gpu_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def fixture_snl1():
    import snl1_fixture
    return snl1_fixture.load()


@pytest.fixture(scope="session")
def gpu_runtime():
    """Initialise the library's HIP runtime once per session (GPU tests only)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ww3gpu.snl1 import init_runtime
    init_runtime()
    return torch
