import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def airfoil_mesh():
    from dgm_in_rust_amd.meshgen import read_gmsh
    return read_gmsh(os.path.join(ROOT, "tests", "golden", "airfoil.msh"))
