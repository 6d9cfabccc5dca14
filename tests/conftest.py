import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "focus-stacker_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """torch carries its own HIP runtime beside the system one that libfocus_gpu.so links.  When
    the library's runtime opens the GPU first, torch's later finds no GPU; opened in the other
    order both work.  So a session that runs GPU tests lets torch open the device before any test
    loads the library (torch only generates frames for some of the tests)."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    yield


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle64():
    from oracle import Oracle
    return Oracle(fp64=True)
