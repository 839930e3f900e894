"""GPU test of the Nose thermostat (scalev, src/thermostat_m.F90:317-395)
through the C ABI, driven as moldin drives it (src/moldin.F90:828-835),
against the reference build's fixture nose_nvt_1024.

Tolerances as for the velocity-Verlet trajectories: energies 1e-11 (PE) and
1e-10 (KE) relative, positions 1e-13 absolute, velocities and forces 1e-9
relative (norm-wise).
"""
import numpy as np
import pytest

from conftest import golden_system, load_golden, run_kwargs
from moldy_amd import engine, initial

pytestmark = pytest.mark.gpu


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / (den if den > 0 else 1.0))


def test_gpu_nose_thermostat_matches_reference():
    F = load_golden("nose_nvt_1024")
    print("[gpu-test] nose thermostat", flush=True)
    eng = engine.Engine(golden_system(F), **run_kwargs(F))
    eng.linkup()
    eng.md_init()
    n = eng.n
    rqke = 1.5 * initial.BK * float(F["run_temprq"]) * (n - 3)
    fs, tt, dt = float(np.float32(0.098227)), 0.0, eng.params.deltat
    tr = []
    for _ in range(int(F["run_nsteps"])):
        tr.append(eng.md_steps(1)[0])
        tt = tt + dt
        if tt >= fs:
            tt = tt - fs
            eng.scalev(float(F["run_nose"]), rqke)
    tr, ref = np.array(tr), F["trace"]
    assert np.max(np.abs(tr[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])) < 1e-11
    assert np.max(np.abs(tr[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])) < 1e-10
    for nm in ("x0", "y0", "z0"):
        assert np.max(np.abs(eng.array(nm) - F[nm + "_f"])) < 1e-13, nm
    v = np.stack([eng.array(k) for k in ("x1", "y1", "z1")], axis=1)
    assert rel(v, np.stack([F[k + "_f"] for k in ("x1", "y1", "z1")], axis=1)) < 1e-9
    f = np.stack([eng.array(k) for k in ("fx", "fy", "fz")], axis=1)
    assert rel(f, np.stack([F[k + "_f"] for k in ("fx", "fy", "fz")], axis=1)) < 1e-9
    assert eng.snhv != 0.0
    eng.close()
