import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def tracer():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blackholerender_amd as bhs
    t = bhs.Tracer(0)
    yield t
    t.close()
