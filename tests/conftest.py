import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def core_lib():
    """libumcp_core.so, built in-tree if it is missing."""
    from umcp_amd import surface
    if not os.path.exists(surface.LIB_PATH):
        from umcp_amd import build
        build.build()
    return surface.load_library()


@pytest.fixture(scope="session")
def mesh_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("meshes")
    os.environ.setdefault("UMCP_CACHE", str(d))
    return d
