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


def test_gpu_new_run_velocities_open_slab_matches_reference():
    """A new run of an open slab (ivol=2) with immobile atoms: velocities from
    the product's setvel and setvel_box (RANMAR after read_system's draws),
    scalev, then the Nose thermostat during the run, against the reference
    build's run of the same (fixture setvel_slab_ivol2_1024, driver flag 2)."""
    from moldy_amd import lattice
    F = load_golden("setvel_slab_ivol2_1024")
    print("[gpu-test] new-run velocities, open slab", flush=True)
    sd = golden_system(F)
    temprq, ivol, boxmass = float(F["run_temprq"]), int(F["run_ivol"]), float(F["run_boxmass"])
    rng = initial.Ranmar(97, 33, 21, 82)
    for _ in range(3 * sd.n):
        rng.rand1()
    dt = sd.deltat_fs * lattice.FS_TO_TIME
    x1 = initial.setvel(rng, sd.z, sd.mass, sd.b0, dt, initial.rqke_init(temprq, sd.n))
    b1 = initial.setvel_box(rng, ivol, sd.n, temprq, dt * dt / (2.0 * sd.n * boxmass))
    eng = engine.Engine(lattice.System(b0=sd.b0, s=sd.s, x1=x1, mass=sd.mass, z=sd.z, deltat_fs=sd.deltat_fs),
                        **run_kwargs(F))
    eng.set_tensor("b1", b1)
    eng.scalev(float(F["run_nose"]), initial.rqke_init(temprq, sd.n))
    eng.linkup()
    eng.md_init()
    rqke = 1.5 * initial.BK * temprq * (sd.n - 3)
    fs, tt = float(np.float32(0.098227)), 0.0
    tr = []
    for _ in range(int(F["run_nsteps"])):
        tr.append(eng.md_steps(1)[0])
        tt = tt + eng.params.deltat
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
    np.testing.assert_array_equal(eng.tensor("b1"), F["b1_f"])
    np.testing.assert_array_equal(eng.tensor("b0"), F["b0_f"])
    eng.close()
