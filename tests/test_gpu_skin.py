"""GPU tests of the device list's skin bound and of list rebuilds inside the
MD loop.

The device list holds pairs within r_int + skin.  A step may use it while
2 d + e r_list < skin, d the largest atom displacement and e the box strain
(Frobenius norm of B0 Bref^-1 - I) since the list was built; the engine
rebuilds before that can fail and reports MDK_ERR_SKIN when a completed step
shows it failed.  Trajectory tolerances are the FP64 parity bars (energy 1e-9
relative, fractional positions 1e-11).
"""
import sys

import numpy as np
import pytest

from moldy_amd import engine, lattice

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def test_gpu_skin_bound_violation_reported():
    """An atom that moves 0.2 A per step outruns the 0.15 A half-skin before the
    monitor (two steps late) can ask for a rebuild: md_steps reports
    MDK_ERR_SKIN instead of carrying on with a list that may miss pairs."""
    sd = lattice.make_system(nc=8, sigma=0.05, temp=300.0, seed=2)
    say("skin violation")
    eng = engine.Engine(sd, ivol=1, ntcm=1000, rebuild=1)
    eng.linkup()
    eng.md_init()
    x1 = eng.array("x1")
    x1[17] = 0.2 / sd.b0[0, 0]  # fractional displacement per step along x
    eng.upload("x1", x1)
    with pytest.raises(engine.MdkError) as ei:
        eng.md_steps(8)
    assert ei.value.code == engine.MDK_ERR_SKIN
    eng.close()


def test_gpu_box_strain_triggers_rebuilds(oracle):
    """Parrinello-Rahman NPT at 40 GPa from a cold crystal: the box shrinks by
    10 % in 120 steps while atoms barely move, so the strain terms drive the
    rebuilds; the run matches a reference run that rebuilds its list every
    step."""
    sd = lattice.make_system(nc=8, sigma=0.02, temp=10.0, seed=5)
    say("strain-driven rebuilds")
    kw = dict(ivol=0, press_gpa=40.0, boxmass=5.0)
    eng = engine.Engine(sd, ntcm=1000, rebuild=1, **kw)
    eng.linkup()
    eng.md_init()
    tr = eng.md_steps(120)
    rebuilds, _ = eng.stats()
    S = oracle.OracleSystem(sd, oracle.Potential(), ntcm=1, **kw)
    S.call("linkup")
    S.call("md_init")
    for _ in range(120):
        S.call("md_step")
    strain = np.linalg.norm(eng.tensor("b0") @ np.linalg.inv(sd.b0) - np.eye(3))
    say(f"box strain {strain:.4f}, rebuilds {rebuilds}")
    assert strain > 0.08
    assert rebuilds >= 2
    assert abs(tr[-1, 0] - S.scalar("pe")) / abs(S.scalar("pe")) < 1e-9
    assert max(np.max(np.abs(eng.array(f) - S.array(f))) for f in ("x0", "y0", "z0")) < 1e-11
    assert np.max(np.abs(eng.tensor("b0") - S.tensor("b0"))) < 1e-9
    eng.close()


def test_gpu_strain_loop_list_overflow_recovers(oracle):
    """A list capacity that the compressed box overflows: a strain loop that
    compresses the box by 2 % between two runs makes the next list inside the
    MD loop overflow; that build is discarded, the capacity grown and the list
    built again before any force uses it, so the run continues and matches the
    reference."""
    sd = lattice.make_system(nc=8, sigma=0.05, temp=300.0, seed=9)
    say("strain loop with list overflow")
    eng = engine.Engine(sd, ivol=1, ntcm=10, max_full=24)
    eng.linkup()
    eng.md_init()
    S = oracle.OracleSystem(sd, oracle.Potential(), ivol=1, ntcm=10)
    S.call("linkup")
    S.call("md_init")
    eng.md_steps(15)
    for _ in range(15):
        S.call("md_step")
    b0 = eng.tensor("b0") * 0.98
    eng.set_tensor("b0", b0)
    S.set_tensor("b0", b0)
    tr = eng.md_steps(15)
    for _ in range(15):
        S.call("md_step")
    # the compressed crystal has atoms with more than 24 list partners
    from scipy.spatial import cKDTree
    L = b0[0, 0]
    pos = (np.stack([eng.array(f) for f in ("x0", "y0", "z0")], axis=1) % 1.0) * L
    rl = 3.72645 + 0.3
    most = max(len(x) - 1 for x in cKDTree(pos, boxsize=L).query_ball_point(pos, rl))
    say(f"largest list row after compression {most}")
    assert most > 24
    assert abs(tr[-1, 0] - S.scalar("pe")) / abs(S.scalar("pe")) < 1e-9
    assert max(np.max(np.abs(eng.array(f) - S.array(f))) for f in ("x0", "y0", "z0")) < 1e-11
    eng.close()
