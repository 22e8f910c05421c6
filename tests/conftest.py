"""Shared test setup: the `gpu` marker and import paths.

The oracle (oracle/) is the checker: tests compare the HIP product against
it, never the other way round.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "matlab-cuda-efief-vs-cpu_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def mirror():
    import mirror_np
    return mirror_np


@pytest.fixture(scope="session")
def efie_mod():
    import efie
    efie.lib()
    return efie
