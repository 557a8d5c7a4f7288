"""Shared pytest setup: marker registration and import paths.

`-m "not gpu"` runs the oracle checks, the host-side logic and the C-ABI
load/export checks; `-m gpu` runs the HIP path against the oracle.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "dav-qed2d_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def O():
    from oracle import qed2d_oracle
    return qed2d_oracle


@pytest.fixture(scope="session")
def Q():
    """The HIP library's Python mirror, loaded on the GPU."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qed2d
    return qed2d
