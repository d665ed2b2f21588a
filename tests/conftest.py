import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libcrvpt.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.load()
    return O


@pytest.fixture(scope="module")
def vpt():
    """The gfx950 renderer, initialised once per test module on cuda:0."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cloudrendering_amd import vpt as V

    V.initialize()
    yield V
    V.cleanup()
