import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "warp-and-project-tomography_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    import pyoracle
    if not pyoracle.available_ref():
        pytest.skip("reference-compiled checker (oracle/_ref) not built here")
    return pyoracle.Ref()


@pytest.fixture(scope="session")
def wctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import wap_amd
    ctx = wap_amd.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def W():
    """The package's Python mirror of the C ABI."""
    import wap_amd
    return wap_amd
