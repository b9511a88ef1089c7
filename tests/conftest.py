import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "lammps-lmff_amd"), os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

# The library keeps systems below 100,000 atoms in the pull ILP kernel; the GPU tests run their
# (small) systems through the half-pair kernel wherever it applies, as the large ones run, and
# compare the two kernels explicitly (LMFF_HALF=0 selects the pull kernel).  Set before the
# library reads it at its first context.
os.environ.setdefault("LMFF_HALF_MIN_ATOMS", "0")

POT = os.path.join(REPO, "lammps-lmff_amd", "potentials")
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle
