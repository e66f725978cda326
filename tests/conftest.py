import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


HAVE_GPU = None


def pytest_collection_modifyitems(config, items):
    global HAVE_GPU
    for item in items:
        if "gpu" in item.keywords:
            if HAVE_GPU is None:
                HAVE_GPU = _have_gpu()
            if not HAVE_GPU:
                item.add_marker(pytest.mark.skip(reason="no HIP device"))


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.build(ref=os.path.isdir("/root/reference/src"))
    return oracle
