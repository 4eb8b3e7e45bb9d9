import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def ref():
    """The oracle (CPU restatement of NT2's scalar path)."""
    import nt2ref

    nt2ref.lib()
    return nt2ref


@pytest.fixture(scope="session")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu test without a visible HIP device")
    import nt2_amd

    nt2_amd.call("nt2h_init", 0)
    return torch.device("cuda:0")
