"""Test configuration.

`gpu` marks tests that need an MI355X; the CPU suite runs with -m "not gpu".
Native libraries are built in-tree by __graft_entry__.build(); the session
builds them here only when they are missing.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)

_LIBS = [os.path.join(ROOT, "oracle", "build", "libocl_oracle.so"),
         os.path.join(ROOT, "opencamlib_amd", "lib", "libocl_hip.so"),
         os.path.join(ROOT, "opencamlib_amd", "lib", "libocl_selcheck.so")]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: full-size cases")
    if not all(os.path.exists(p) for p in _LIBS):
        from opencamlib_amd import build as B
        B.build_all()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    torch.cuda.set_device(0)
    return torch.device("cuda:0")
