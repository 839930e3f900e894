"""CPU tests of the C ABI boundary and host-side logic (no GPU compute calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from moldy_amd import engine, lattice

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "moldy_hip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(engine.LIB_PATH)
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(engine.EXPORTS) == sorted(n for n in names if n in engine.EXPORTS)
    assert set(names) == set(engine.EXPORTS) | set(getattr(engine, "EXPORTS_EXTRA", []))


def test_params_struct_matches_header():
    text = open(HEADER).read()
    body = re.search(r"typedef struct mdk_params \{(.*?)\} mdk_params;", text, re.S).group(1)
    fields = re.findall(r"\b(?:int|double)\s+([a-z0-9_, ]+);", body)
    names = [f.strip() for grp in fields for f in grp.split(",")]
    assert names == [f[0] for f in engine.MdkParams._fields_]


def test_create_without_gpu_reports_no_device():
    lib = engine.load_library()
    p = engine.MdkParams(nm=10, ivol=1, nnbrs=150, ntcm=10, nlcx=3, nlcy=3, nlcz=3,
                         precision=0, rebuild=0, max_full=0, deltat=0.1, rnear=6.3, rskin=0.3,
                         press=0.0, bdel2=1.0)
    h = C.c_void_p()
    rc = lib.mdk_create(C.byref(h), C.byref(p))
    if rc == engine.MDK_OK:
        lib.mdk_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc == engine.MDK_ERR_NO_DEVICE
    assert b"device" in lib.mdk_last_error()


def test_bad_parameters_rejected():
    lib = engine.load_library()
    p = engine.MdkParams(nm=0, ivol=1, nnbrs=150, ntcm=10, nlcx=3, nlcy=3, nlcz=3, deltat=0.1, rnear=6.3)
    h = C.c_void_p()
    assert lib.mdk_create(C.byref(h), C.byref(p)) == engine.MDK_ERR_BAD_ARG
    p.nm, p.nlcx = 10, 2  # moldin forces at least 3 link cells per dimension
    assert lib.mdk_create(C.byref(h), C.byref(p)) == engine.MDK_ERR_BAD_ARG


def test_link_cell_counts_follow_moldin():
    b0 = np.diag([28.665, 17.199, 183.456])
    assert engine.link_cells_for(b0, 6.3) == (4, 3, 29)


def test_lattice_generator():
    sd = lattice.make_system(nc=(4, 5, 6), a=2.8665, sigma=0.05, temp=300.0, seed=2)
    assert sd.n == 2 * 4 * 5 * 6
    v = sd.x1 @ sd.b0.T / (1.0 * lattice.FS_TO_TIME)
    p = (v * sd.mass[:, None]).sum(axis=0)
    assert np.all(np.abs(p) < 1e-9)
    ke = 0.5 * (sd.mass[:, None] * v * v).sum()
    assert abs(ke / (1.5 * sd.n * lattice.BK) - 300.0) < 1e-9
    sv = lattice.make_system(nc=5, vacancy_frac=0.02, seed=3)
    assert sv.n == 250 - 5


def test_liquid_lattice_parameter():
    assert abs(lattice.liquid_lattice_parameter(7.0) - 2.98118) < 1e-4
