import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmoldy_hip.so)")


def load_golden(name):
    """Golden fixture written by tests/golden/gen_golden.py (no pickles)."""
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def golden_system(F):
    from moldy_amd import lattice
    return lattice.System(b0=F["b0"], s=F["s"], x1=F["x1"], mass=F["mass"], z=F["z"],
                          deltat_fs=float(F["deltat_fs"]))


def run_kwargs(F):
    kw = dict(ivol=int(F["run_ivol"]), ntcm=int(F["run_ntcm"]))
    if "run_press" in F:
        kw["press_gpa"] = float(F["run_press"])
    if "run_boxmass" in F:
        kw["boxmass"] = float(F["run_boxmass"])
    return kw


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle
