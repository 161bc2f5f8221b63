import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "latticeboltzmann.jl_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and liblbm_hip.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _gpu_available() -> bool:
    # a library that is missing or lacks a declared symbol raises here instead
    # of letting every GPU test skip silently; only "no device" skips
    import lbmhip
    return lbmhip.device_count() > 0


def pytest_collection_modifyitems(config, items):
    if not any(item.get_closest_marker("gpu") for item in items):
        return
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if item.get_closest_marker("gpu"):
            item.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import clib
    clib.build()
    return clib
