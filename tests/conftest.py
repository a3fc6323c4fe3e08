import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def _ensure_built():
    lib = os.path.join(ROOT, "controllerengine_amd", "libControllerEngine.so")
    ora = os.path.join(ROOT, "oracle", "libce_oracle.so")
    if not os.path.exists(ora):
        from oracle import build as ob
        ob.build(verbose=False)
    if not os.path.exists(lib):
        from controllerengine_amd import build as pb
        pb.build(verbose=False)


_ensure_built()


@pytest.fixture(scope="session")
def engine():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from controllerengine_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()
