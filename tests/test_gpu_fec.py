"""GPU tests of the iron_carbon material (src/iron_carbon.F90) through the C ABI.

Against the reference's iron_carbon build (fixtures tests/golden/fec_*.npz)
and against the reference's own shipped test tests/reference/bcc_iron
(tests/golden/bcc_iron_golden.npz: what output66.txt and system.out print).

Tolerances:
  * lookup-mode potential values: bit-exact (host-built tables, the
    reference's index arithmetic);
  * exact functions: 1e-13 relative above 2.05 A, 1e-12 below (exp on the
    device); embedding 1e-14 relative with a 1e-14 absolute floor;
  * FP64 trajectories: as for fetransmetal (forces 1e-10 / 1e-9 relative,
    energies 1e-11, volume and box 1e-12);
  * bcc_iron: every printed digit of output66.txt's averages, temperatures,
    virial stress diagonal and box diagonal, and of system.out's sampled
    positions and energies; the box's and stress's off-diagonal elements
    (1e-8 A and 1e-3 GPa: differences of large sums that cancel) within
    1e-6 relative.
"""
import sys

import numpy as np
import pytest

from conftest import golden_system, load_golden, run_kwargs
from moldy_amd import engine

pytestmark = pytest.mark.gpu

CASES = ["fec_npt_1024", "fec_lookup_2000", "fec_hot_1024"]


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def make_engine(F, **kw):
    args = run_kwargs(F)
    args.update(kw)
    if int(F["run_uselookup"]):
        args.setdefault("lookup", 5000)
    return engine.Engine(golden_system(F), material="iron_carbon", **args)


def test_gpu_fec_functions_exact():
    F = load_golden("fec_npt_1024")
    say("iron_carbon functions, exact mode")
    eng = make_engine(F)
    r = F["rgrid"]
    for fn in ("vee", "dvee", "phi", "dphi"):
        g, ref = eng.eval_potential(fn, r), F[fn]
        knot = r > float(np.float32(2.05))
        assert rel(g[knot], ref[knot]) < 1e-13, fn
        err = np.abs(g[~knot] - ref[~knot]) / np.maximum(np.abs(ref[~knot]), 1e-300)
        assert np.max(err) < 1e-12, (fn, np.max(err))
        assert np.all(g[ref == 0.0] == 0.0), fn
    for fn in ("emb", "demb"):
        # F' has a root near rho = 28: an absolute floor at the terms' rounding
        np.testing.assert_allclose(eng.eval_potential(fn, F["rhogrid"]), F[fn], rtol=1e-14, atol=1e-14)
    with pytest.raises(engine.MdkError):   # set_up constants exist for fetransmetal only
        eng.potential_constants()
    eng.close()


def test_gpu_fec_lookup_functions_bit_exact():
    F = load_golden("fec_lookup_2000")
    say("iron_carbon functions, lookup mode")
    eng = make_engine(F)
    for fn in ("vee", "dvee", "phi", "dphi"):
        np.testing.assert_array_equal(eng.eval_potential(fn, F["rgrid"]), F[fn], err_msg=fn)
    for fn in ("emb", "demb"):
        np.testing.assert_array_equal(eng.eval_potential(fn, F["rhogrid"]), F[fn], err_msg=fn)
    eng.close()


@pytest.mark.parametrize("case", CASES)
def test_gpu_fec_trajectory_matches_reference(case):
    F = load_golden(case)
    say("iron_carbon trajectory", case)
    eng = make_engine(F)
    eng.linkup()
    eng.md_init()
    for nm, key in [("rho", "rho0"), ("dafrho", "dafrho0"), ("afrho", "afrho0")]:
        assert rel(eng.array(nm), F[key]) < 1e-10, nm
    f = np.stack([eng.array("fx"), eng.array("fy"), eng.array("fz")], axis=1)
    assert rel(f, np.stack([F["fx0"], F["fy0"], F["fz0"]], axis=1)) < 1e-10
    for nm in ("tp", "tk", "fb"):
        assert rel(eng.tensor(nm), F[nm + "0"]) < 1e-10, nm
    tr = eng.md_steps(int(F["run_nsteps"]))
    ref = F["trace"]
    assert np.max(np.abs(tr[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])) < 1e-11   # PE
    assert np.max(np.abs(tr[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])) < 1e-10   # KE
    assert np.max(np.abs(tr[:, 5] - ref[:, 4]) / np.abs(ref[:, 4])) < 1e-13   # volume
    for nm in ("x0", "y0", "z0"):
        assert np.max(np.abs(eng.array(nm) - F[nm + "_f"])) < 1e-13, nm
    v = np.stack([eng.array(n) for n in ("x1", "y1", "z1")], axis=1)
    assert rel(v, np.stack([F[n + "_f"] for n in ("x1", "y1", "z1")], axis=1)) < 1e-10
    f = np.stack([eng.array(n) for n in ("fx", "fy", "fz")], axis=1)
    assert rel(f, np.stack([F[n + "_f"] for n in ("fx", "fy", "fz")], axis=1)) < 1e-9
    for nm in ("b0", "b1"):
        assert np.max(np.abs(eng.tensor(nm) - F[nm + "_f"])) <= 1e-12 * np.max(np.abs(F["b0_f"])), nm
    eng.close()


@pytest.mark.parametrize("lookup", [0, 5000])
def test_gpu_energy_calc_matches_oracle(oracle, lookup):
    """energy_calc: per-atom energies F(rho_i) + 1/2 sum_j V(r_ij) with the exact V."""
    F = load_golden("fec_npt_1024")
    say("energy_calc, lookup", lookup)
    eng = make_engine(F, lookup=lookup)
    eng.linkup()
    eng.md_init()
    en = eng.energy_calc()
    S = oracle.OracleSystem(golden_system(F), oracle.Potential(material="iron_carbon", lookup=bool(lookup)),
                            **run_kwargs(F))
    S.call("linkup")
    S.call("md_init")
    ref = S.energy_calc()
    assert rel(en, ref) < 1e-12
    eng.close()


def test_gpu_fec_refuses_fp32_and_carbon():
    F = load_golden("fec_npt_1024")
    with pytest.raises(engine.MdkError):
        make_engine(F, precision=engine.PREC_FP32)
    sd = golden_system(F)
    sd.z = sd.z.copy()
    sd.z[3] = 6
    with pytest.raises(engine.MdkError):
        engine.Engine(sd, material="iron_carbon", **run_kwargs(F))


def test_gpu_bcc_iron_reference_test(oracle):
    """The reference's shipped test tests/reference/bcc_iron on the GPU path:
    31,250 atoms, iron_carbon lookup tables, PR NPT at 0.1 GPa, 10 steps from
    the state moldin builds; output66.txt and system.out as printed."""
    import bcc_iron_case as bc
    F = bc.load()
    sd = bc.initial_state(F)
    say("bcc_iron, n =", sd.n)
    eng = engine.Engine(sd, material="iron_carbon", lookup=5000, ivol=int(F["ivol"]), rpad=float(F["rpad"]),
                        nnbrs=int(F["nnbrs"]), ntcm=int(F["ntcm"]), press_gpa=float(F["press_gpa"]),
                        boxmass=float(F["boxmass"]), skin=0.3, rebuild=1)
    eng.linkup()
    eng.md_init()
    steps = []
    for _ in range(int(F["nsteps"])):
        tr = eng.md_steps(1)[0]     # pe, ke, te, h, th, vol
        steps.append((tr[0], tr[1], tr[5], tr[4] - tr[3], eng.tensor("b0")))
    got = bc.summarize(steps, eng.tensor("tp"), eng.tensor("b0"), eng.params.press, sd.n, oracle.bk())
    bad = bc.check(F, got, offdiag_rtol=1e-6)
    say("bcc_iron PE %.10e KE %.10e T %.6f" % (got["pe"], got["ke"], got["t"]))
    assert bad == []
    en = eng.energy_calc()
    bad = bc.check_system_out(F, eng.tensor("b0"), eng.array("x0"), eng.array("y0"), eng.array("z0"), en,
                              box_rtol=1e-13, offdiag_rtol=1e-6)
    assert bad == []
    eng.close()
