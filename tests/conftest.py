"""Shared set-up of the test suite.

`gpu`-marked tests need an MI355X (they run the HIP library); everything
else runs on the CPU: the oracle against closed forms and fixtures, the
host-side geometry helpers and the C-ABI symbol table.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "latticeboltzmannmethod-fonnesu-grassi-guffanti_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def lbm():
    import lbm_amd
    lbm_amd.lib()
    return lbm_amd
