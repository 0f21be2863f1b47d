import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    lib = os.path.join(ROOT, "vship_amd", "libvship_amd.so")
    olib = os.path.join(ROOT, "oracle", "liboracle.so")
    mlib = os.path.join(ROOT, "tests", "vs_host", "libmock_vs.so")
    drv = os.path.join(ROOT, "vship_amd", "ffvship_amd")
    if not all(os.path.exists(p) for p in (lib, olib, mlib, drv)):
        import __graft_entry__
        __graft_entry__.build()


@pytest.fixture(scope="session")
def ref_lib_path():
    p = os.path.join(ROOT, "oracle", "_ref", "libvship_ref.so")
    return p if os.path.exists(p) else None
