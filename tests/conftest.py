"""Shared fixtures.  GPU tests are marked `gpu`; they skip only when the host
has no GPU at all (a missing engine library is an error, never a skip)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "times-c_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _ensure_built():
    lib = os.path.join(ROOT, "times-c_amd", "lib", "libtimesc_hip.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if not os.path.exists(lib):
        import __graft_entry__

        __graft_entry__.build()


_ensure_built()


def gpu_present():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def ctx():
    if not gpu_present():
        pytest.skip("no GPU on this host")
    import timesc

    c = timesc.Context(0)
    yield c
    c.close()


def pytest_collection_modifyitems(config, items):
    if gpu_present():
        return
    skip = pytest.mark.skip(reason="no GPU on this host")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
