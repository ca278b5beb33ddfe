import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "pyfacc-scf_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

DATA = os.path.join(ROOT, "pyfacc-scf_amd", "data")
GOLDEN = os.path.join(ROOT, "tests", "golden")
ANG2BOHR = 1.8897259886


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


def load_xyz(name):
    """(species, coords_bohr (natoms,3)) from pyfacc-scf_amd/data/molecules/<name>."""
    import numpy as np
    with open(os.path.join(DATA, "molecules", name)) as fh:
        lines = fh.read().splitlines()
    n = int(lines[0].split()[0])
    sp, xyz = [], []
    for ln in lines[2:2 + n]:
        t = ln.split()
        sp.append(t[0])
        xyz.append([float(t[1]), float(t[2]), float(t[3])])
    return sp, np.array(xyz) * ANG2BOHR


@pytest.fixture(scope="session")
def gpu_lib():
    import pyfacc
    lib = pyfacc.load_library()
    if pyfacc.device_count() < 1:
        pytest.fail("GPU test requested but no GPU is visible")
    return lib
