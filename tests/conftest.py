import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def corpus():
    import reference_corpus as refcorp
    return refcorp.fuzz_corpus()


@pytest.fixture(scope="session")
def gpu():
    """one decoder context for the whole GPU session (it fails loudly without the library)"""
    from gpuzip_amd.gpuinflate import Gpu
    g = Gpu()
    yield g
    g.close()
