"""GPU tests of the Gear predictor-corrector (iverlet = 0, src/quench_m.F90:636-759)
through the C ABI, against the reference build's fixtures gear_*.npz.

Tolerances as for the velocity-Verlet trajectories (test_gpu_parity):
positions 1e-13 absolute (fractional), velocities/forces and the higher
derivatives 1e-9 relative (norm-wise), energies 1e-11, volume 1e-13, box
1e-12 of its largest element.
"""
import sys

import numpy as np
import pytest

from conftest import golden_system, load_golden, run_kwargs
from moldy_amd import engine

pytestmark = pytest.mark.gpu

CASES = ["gear_npt_1024", "gear_z_1024"]


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def make_engine(F, **kw):
    return engine.Engine(golden_system(F), iverlet=0, **run_kwargs(F), **kw)


def check_final(eng, F):
    for nm in ("x0", "y0", "z0"):
        assert np.max(np.abs(eng.array(nm) - F[nm + "_f"])) < 1e-13, nm
    for grp in (("x1", "y1", "z1"), ("x2", "y2", "z2"), ("x3", "y3", "z3"), ("fx", "fy", "fz")):
        g = np.stack([eng.array(n) for n in grp], axis=1)
        r = np.stack([F[n + "_f"] for n in grp], axis=1)
        assert rel(g, r) < 1e-9, grp
    scale = np.max(np.abs(F["b0_f"]))
    for nm in ("b0", "b1", "b2", "b3"):
        assert np.max(np.abs(eng.tensor(nm) - F[nm + "_f"])) <= 1e-12 * scale, nm


@pytest.mark.parametrize("case", CASES)
def test_gpu_gear_trajectory_matches_reference(case):
    F = load_golden(case)
    say("gear trajectory", case)
    eng = make_engine(F)
    eng.linkup()
    eng.md_init()
    tr = eng.md_steps(int(F["run_nsteps"]))
    ref = F["trace"]
    assert np.max(np.abs(tr[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])) < 1e-11   # PE
    assert np.max(np.abs(tr[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])) < 1e-10   # KE
    assert np.max(np.abs(tr[:, 5] - ref[:, 4]) / np.abs(ref[:, 4])) < 1e-13   # volume
    check_final(eng, F)
    eng.close()


def test_gpu_gear_step_api_matches_md_steps():
    """predic, force, correc, update_therm_avs, relink (moldin's call order)
    give the same state as md_steps with iverlet = 0."""
    F = load_golden("gear_npt_1024")
    say("gear step api vs md_steps")
    a, b = make_engine(F), make_engine(F)
    for e in (a, b):
        e.linkup()
        e.md_init()
    tr = a.md_steps(6)
    for k in range(6):
        b.predic()
        b.force()
        b.correc()
        b.update_therm_avs()
        if b.ntc == 0:
            b.relink()
        assert abs(b.scalar("ke") - tr[k, 1]) <= 1e-12 * tr[k, 1]
    for nm in ("x0", "x1", "x2", "x3", "fx"):
        np.testing.assert_array_equal(a.array(nm), b.array(nm), err_msg=nm)
    for nm in ("b0", "b3"):
        np.testing.assert_array_equal(a.tensor(nm), b.tensor(nm), err_msg=nm)
    a.close()
    b.close()


def test_gpu_first_predic_is_inert():
    """Like the reference, predic changes nothing before a correc or
    velocityverlet has run on the handle."""
    F = load_golden("gear_npt_1024")
    eng = make_engine(F)
    eng.linkup()
    eng.md_init()
    x0, x1, b0 = eng.array("x0"), eng.array("x1"), eng.tensor("b0")
    eng.predic()
    np.testing.assert_array_equal(eng.array("x0"), x0)
    np.testing.assert_array_equal(eng.array("x1"), x1)
    np.testing.assert_array_equal(eng.tensor("b0"), b0)
    eng.force()
    eng.correc()
    x0c = eng.array("x0")
    eng.predic()             # now live: x0 += x1 + x2 + x3
    assert not np.array_equal(eng.array("x0"), x0c)
    eng.close()
