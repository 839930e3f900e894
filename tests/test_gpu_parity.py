"""GPU parity tests: libmoldy_hip.so (through its C ABI) against the oracle and
the reference's golden vectors.

Tolerances (stated per quantity):
  * neighbour sets: bit-exact as unordered integer pair sets;
  * FP64 form: forces, densities, F', virial within 1e-10 relative (norm-wise:
    max|a-b| / max|b|); energies within 1e-12 relative;
  * lookup-mode potential values: bit-exact;
  * FP32 form: max force error <= 1e-5 eV/A against the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from conftest import golden_system, load_golden, run_kwargs
from moldy_amd import engine, lattice

pytestmark = pytest.mark.gpu

# open boundaries (ivol 2, 5, 3): the slab has an immobile layer (mass 0)
OPEN_CASES = ["slab_ivol2_1024", "pillar_ivol5_1024", "cluster_ivol3_1024"]
CASES = ["c0_nve_2000", "npt_tric_1024", "hot_1024"] + OPEN_CASES


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
    return engine.Engine(golden_system(F), **args)


def test_gpu_potential_constants_and_functions(oracle):
    F = load_golden("c0_nve_2000")
    say("potential functions, exact mode")
    eng = make_engine(F)
    pot = oracle.Potential()
    c = eng.potential_constants()
    np.testing.assert_array_equal(c[2:6], pot.spline_coeff)
    assert (c[6], c[7]) == pot.phi_cut
    r = F["rgrid"]
    for fn in ("vee", "dvee", "phi", "dphi"):
        g = eng.eval_potential(fn, r)
        ref = F[fn]
        knot = r > float(np.float32(1.9))
        assert rel(g[knot], ref[knot]) < 1e-13, fn
        short = ~knot
        err = np.abs(g[short] - ref[short]) / np.maximum(np.abs(ref[short]), 1e-300)
        assert np.max(err) < 1e-11, (fn, np.max(err))
        assert np.all(g[ref == 0.0] == 0.0), fn
    for fn in ("emb", "demb"):
        g = eng.eval_potential(fn, F["rhogrid"])
        np.testing.assert_allclose(g, F[fn], rtol=1e-15, atol=0)
    eng.close()


def test_gpu_lookup_functions_bit_exact():
    F = load_golden("lookup_2000")
    say("potential functions, lookup mode")
    eng = make_engine(F)
    for fn in ("vee", "dvee", "phi", "dphi"):
        np.testing.assert_array_equal(eng.eval_potential(fn, F["rgrid"]), F[fn], err_msg=fn)
    for fn in ("emb", "demb"):
        np.testing.assert_array_equal(eng.eval_potential(fn, F["rhogrid"]), F[fn], err_msg=fn)
    eng.close()


@pytest.mark.parametrize("case", CASES + ["lookup_2000"])
def test_gpu_reference_half_list_bit_exact(oracle, case):
    F = load_golden(case)
    say("half list", case)
    eng = make_engine(F)
    eng.linkup()
    eng.update_neighbourlist()
    pairs = eng.pairs()
    assert len(pairs) == int(F["npairs0"])
    assert oracle.pairset_digest(pairs) == str(F["pairs0_sha256"])
    numn, nlist = eng.neighbourlist()
    assert int(numn.sum()) == int(F["npairs0"])
    # link cells as linkup leaves them
    S = oracle.OracleSystem(golden_system(F), oracle.Potential(), **run_kwargs(F))
    S.call("linkup")
    l0, link, ic = eng.linkcells()
    np.testing.assert_array_equal(ic, S.ic())
    for c in range(len(l0)):  # closed chains visiting exactly the cell's atoms, ascending
        if l0[c] == 0:
            continue
        members, i = [], l0[c]
        while True:
            members.append(i)
            i = link[i - 1]
            if i == l0[c]:
                break
        assert members == sorted(members)
        assert np.all(ic[np.array(members) - 1] == c + 1)
    eng.close()


def test_gpu_relink_chains_and_half_list_match_reference(oracle):
    """relink keeps the reference's chain order (src/linkcell_m.F90:224-271:
    an atom that changes cell becomes the new cell's starter, the atom after
    it the old cell's), and the half list follows it: which atom of an
    own-cell pair lists the other, and the order inside NLIST(:, i).  After
    two rounds of random displacements and relinks, L0, LINK and IC equal the
    oracle's, and so does every NLIST(:, i), entry for entry."""
    F = load_golden("hot_1024")
    say("relink chains")
    eng = make_engine(F)
    S = oracle.OracleSystem(golden_system(F), oracle.Potential(), **run_kwargs(F))
    eng.linkup()
    S.call("linkup")
    rng = np.random.default_rng(5)
    s = golden_system(F).s.copy()
    moved = 0
    for _ in range(2):
        s = s + rng.normal(0.0, 0.03, s.shape)
        for k, nm in enumerate(("x0", "y0", "z0")):
            eng.upload(nm, s[:, k])
            S.set_array(nm, s[:, k])
        ic_before = S.ic()
        eng.relink()
        S.call("relink")
        moved += int(np.sum(S.ic() != ic_before))
        l0, link, ic = eng.linkcells()
        l0o, linko = S.chains()
        np.testing.assert_array_equal(ic, S.ic())
        np.testing.assert_array_equal(l0, l0o)
        np.testing.assert_array_equal(link, linko)
    assert moved > 100
    eng.update_neighbourlist()
    assert S.call("update_neighbourlist") == 0
    np.testing.assert_array_equal(eng.pairs(), S.pairs())
    eng.close()


def test_gpu_resort_keeps_what_is_read_again():
    """A list rebuild outside force (rhoset on moved positions) re-sorts every
    field that can still be read: the forces of the last force evaluation come
    back unchanged in reference order, and so do third derivatives that were
    never set (zero) after MD steps with velocity Verlet."""
    F = load_golden("npt_tric_1024")
    say("re-sort outside force")
    eng = make_engine(F)
    eng.linkup()
    eng.md_init()
    eng.md_steps(12)  # several rebuilds inside force
    for nm in ("x3", "y3", "z3"):
        assert not np.any(eng.array(nm)), nm
    f = np.stack([eng.array(nm) for nm in ("fx", "fy", "fz")], axis=1)
    x = eng.array("x0")
    eng.upload("x0", x + 1e-3)  # moved positions: the next pass rebuilds the list
    eng.call("rhoset")
    np.testing.assert_array_equal(np.stack([eng.array(nm) for nm in ("fx", "fy", "fz")], axis=1), f)
    np.testing.assert_array_equal(eng.array("x0"), x + 1e-3)
    eng.close()


def test_gpu_atomic_stress_matches_reference():
    """mdk_atomic_stress after the first force and its half list: the
    reference's per-atom diagonal stress (atomstress,
    src/dynamics_m.F90:279-283) on the sheared NPT box to 1e-10 of the largest
    component; before any half list the entry point refuses."""
    from moldy_amd.engine import MdkError
    F = load_golden("atomstress_tric_1024")
    say("atomic stress")
    eng = make_engine(F)
    eng.linkup()
    with pytest.raises(MdkError):
        eng.atomic_stress()
    eng.md_init()
    eng.update_neighbourlist()
    assert rel(eng.atomic_stress(), F["stress0"]) < 1e-10
    eng.close()


@pytest.mark.parametrize("case", CASES + ["lookup_2000"])
def test_gpu_first_force_matches_reference(case):
    F = load_golden(case)
    say("first force", case)
    eng = make_engine(F)
    eng.linkup()
    eng.md_init()
    for nm, key in [("rho", "rho0"), ("dafrho", "dafrho0"), ("afrho", "afrho0")]:
        assert rel(eng.array(nm), F[key]) < 1e-10, nm
    f = np.stack([eng.array("fx"), eng.array("fy"), eng.array("fz")], axis=1)
    fr = np.stack([F["fx0"], F["fy0"], F["fz0"]], axis=1)
    assert rel(f, fr) < 1e-10
    x2 = np.stack([eng.array("x2"), eng.array("y2"), eng.array("z2")], axis=1)
    assert rel(x2, np.stack([F["x20"], F["y20"], F["z20"]], axis=1)) < 1e-10
    s = eng.scalars()
    assert abs(s["pe"] - float(F["pe0"])) < 1e-12 * abs(float(F["pe0"]))
    for nm in ("tp", "tk", "fb"):
        assert rel(eng.tensor(nm), F[nm + "0"]) < 1e-10, nm
    np.testing.assert_array_equal(eng.tensor("tc"), F["tc0"])
    assert rel(eng.tensor("tgid"), F["tgid0"]) < 1e-10 or np.max(np.abs(F["tgid0"])) == 0
    eng.close()


@pytest.mark.parametrize("case", CASES + ["lookup_2000"])
def test_gpu_md_trajectory_matches_reference(case):
    F = load_golden(case)
    say("trajectory", case)
    eng = make_engine(F)
    eng.linkup()
    eng.md_init()
    tr = eng.md_steps(int(F["run_nsteps"]))
    ref = F["trace"]
    assert np.max(np.abs(tr[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])) < 1e-11   # PE
    assert np.max(np.abs(tr[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])) < 1e-10   # KE
    assert np.max(np.abs(tr[:, 5] - ref[:, 4]) / np.abs(ref[:, 4])) < 1e-13   # volume
    for nm in ("x0", "y0", "z0"):
        assert np.max(np.abs(eng.array(nm) - F[nm + "_f"])) < 1e-13, nm
    v = np.stack([eng.array(n) for n in ("x1", "y1", "z1")], axis=1)
    vr = np.stack([F[n + "_f"] for n in ("x1", "y1", "z1")], axis=1)
    assert rel(v, vr) < 1e-10
    f = np.stack([eng.array(n) for n in ("fx", "fy", "fz")], axis=1)
    fr = np.stack([F[n + "_f"] for n in ("fx", "fy", "fz")], axis=1)
    assert rel(f, fr) < 1e-9
    for nm in ("b0", "b1"):
        assert np.max(np.abs(eng.tensor(nm) - F[nm + "_f"])) <= 1e-12 * np.max(np.abs(F["b0_f"])), nm
    assert rel(eng.tensor("tp"), F["tp_f"]) < 1e-9
    eng.close()


def test_gpu_step_api_matches_md_steps():
    """velocityverlet + update_therm_avs + relink (the Fortran call order)
    gives the same state as the fused md_steps path."""
    F = load_golden("npt_tric_1024")
    say("step api vs fused")
    a, b = make_engine(F), make_engine(F)
    for e in (a, b):
        e.linkup()
        e.md_init()
    tr = a.md_steps(6)
    for k in range(6):
        b.velocityverlet()
        s1, s2, s3 = b.update_therm_avs()
        if b.ntc == 0:
            b.relink()
        assert abs(b.scalar("ke") - tr[k, 1]) <= 1e-12 * tr[k, 1]
    say("list builds", a.stats()[0], b.stats()[0])
    assert a.stats()[0] == b.stats()[0]  # the same skin decisions on the same samples
    for nm in ("x0", "x1", "fx"):
        np.testing.assert_array_equal(a.array(nm), b.array(nm))
    a.close()
    b.close()


def test_gpu_fp32_form_force_error(oracle):
    F = load_golden("c0_nve_2000")
    say("fp32 form")
    eng = make_engine(F, precision=engine.PREC_FP32)
    eng.linkup()
    eng.md_init()
    f32 = eng.forces_cart()
    S = oracle.OracleSystem(golden_system(F), oracle.Potential(), **run_kwargs(F))
    S.call("linkup")
    S.call("md_init")
    f64 = S.forces_cart()
    err = np.max(np.abs(f32 - f64))
    say(f"fp32 max force error {err:.3e} eV/A, pe rel err "
        f"{abs(eng.scalar('pe') - S.scalar('pe')) / abs(S.scalar('pe')):.3e}")
    assert err <= 1e-5                                       # eV/A
    assert abs(eng.scalar("pe") - S.scalar("pe")) / S.n < 1e-6  # eV/atom
    # virial as stress (the tp sum cancels ~1e4-fold, so compare in GPa)
    ev_a3_to_gpa = 160.21766208
    dsig = np.max(np.abs(eng.tensor("tp") - S.tensor("tp"))) / S.scalar("vol") * ev_a3_to_gpa
    say(f"fp32 virial error {dsig:.3e} GPa")
    assert dsig < 1e-4
    eng.close()


@pytest.mark.parametrize("case", OPEN_CASES)
def test_gpu_fp32_form_open_boundaries(oracle, case):
    """FP32 form with open axes (atoms straddling the fold at fractional 0):
    first forces within 1e-5 eV/A of the oracle."""
    F = load_golden(case)
    say("fp32 form, open boundaries", case)
    eng = make_engine(F, precision=engine.PREC_FP32)
    eng.linkup()
    eng.md_init()
    S = oracle.OracleSystem(golden_system(F), oracle.Potential(), **run_kwargs(F))
    S.call("linkup")
    S.call("md_init")
    assert np.max(np.abs(eng.forces_cart() - S.forces_cart())) <= 1e-5
    eng.close()


@pytest.mark.parametrize("nc", [3, 5])
def test_gpu_small_box_legacy_list(oracle, nc):
    """Boxes with fewer than four list cells per dimension (2 and 3 here) take
    the legacy full list: absolute packed positions, image-shift codes, and
    per-pair rounding below three cells.  FP64: a 20-step PR NPT trajectory
    against the oracle; FP32 form: the first forces within 1e-5 eV/A."""
    sd = lattice.make_system(nc=nc, sigma=0.08, temp=300.0, seed=40 + nc)
    say("small box", nc)
    S = oracle.OracleSystem(sd, oracle.Potential(), ivol=0, ntcm=10)
    S.call("linkup")
    S.call("md_init")
    f_init = S.forces_cart()
    e32 = engine.Engine(sd, ivol=0, ntcm=10, precision=engine.PREC_FP32)
    e32.linkup()
    e32.md_init()
    assert np.max(np.abs(e32.forces_cart() - f_init)) <= 1e-5  # eV/A
    e32.close()
    eng = engine.Engine(sd, ivol=0, ntcm=10)
    eng.linkup()
    eng.md_init()
    assert rel(eng.forces_cart(), f_init) < 1e-10
    tr = eng.md_steps(20)
    pe = []
    for _ in range(20):
        S.call("md_step")
        pe.append(S.scalar("pe"))
    assert np.max(np.abs(tr[:, 0] - np.array(pe)) / np.abs(np.array(pe))) < 1e-11
    assert rel(eng.forces_cart(), S.forces_cart()) < 1e-9
    assert np.max(np.abs(eng.array("x0") - S.array("x0"))) < 1e-12
    eng.close()


def test_gpu_per_atom_masses(oracle):
    """Masses that differ within a species keep the kick on the per-atom mass
    and del2 fields (the species table only holds when all atoms of a species
    agree): an NVE trajectory against the oracle."""
    sd = lattice.make_system(nc=6, sigma=0.05, temp=300.0, seed=77)
    sd.mass[::7] *= 1.25
    say("per-atom masses")
    eng = engine.Engine(sd, ivol=1, ntcm=10)
    eng.linkup()
    eng.md_init()
    tr = eng.md_steps(20)
    S = oracle.OracleSystem(sd, oracle.Potential(), ivol=1, ntcm=10)
    S.call("linkup")
    S.call("md_init")
    ke = []
    for _ in range(20):
        S.call("md_step")
        ke.append(S.scalar("ke"))
    assert np.max(np.abs(tr[:, 1] - np.array(ke)) / np.abs(np.array(ke))) < 1e-10
    v = np.stack([eng.array(n) for n in ("x1", "y1", "z1")], axis=1)
    vr = np.stack([S.array(n) for n in ("x1", "y1", "z1")], axis=1)
    assert rel(v, vr) < 1e-10
    eng.close()


def test_gpu_neighbour_overflow_reported():
    sd = lattice.make_system(nc=6, sigma=0.05, seed=1)
    say("overflow")
    eng = engine.Engine(sd, nnbrs=10)
    eng.linkup()
    with pytest.raises(engine.MdkError) as ei:
        eng.update_neighbourlist()
    assert ei.value.code == engine.MDK_ERR_NBR_OVERFLOW
    eng.close()


def test_gpu_skin_rebuild_hot_liquid(oracle):
    """Displacement-triggered rebuilds of the pruned full list reproduce a
    reference run that rebuilds its (rnear) list every step."""
    sd = lattice.make_system(nc=6, a=lattice.liquid_lattice_parameter(7.0), sigma=0.15,
                             temp=2500.0, deltat_fs=0.5, seed=21)
    say("skin rebuild")
    eng = engine.Engine(sd, ivol=1, ntcm=1000, rebuild=1)
    eng.linkup()
    eng.md_init()
    tr = eng.md_steps(60)
    rebuilds, _ = eng.stats()
    S = oracle.OracleSystem(sd, oracle.Potential(), ivol=1, ntcm=1)
    S.call("linkup")
    S.call("md_init")
    pe = []
    for _ in range(60):
        S.call("md_step")
        pe.append(S.scalar("pe"))
    assert rebuilds >= 2
    assert np.max(np.abs(tr[:, 0] - np.array(pe)) / np.abs(np.array(pe))) < 1e-9
    x = eng.array("x0")
    assert np.max(np.abs(x - S.array("x0"))) < 1e-11
    eng.close()


def test_gpu_large_perfect_lattice_closed_form(oracle):
    """64^3 bcc cells (524,288 atoms): perfect-lattice energy per atom and zero
    forces against the closed form."""
    a = 2.8665
    sd = lattice.make_system(nc=64, a=a, sigma=0.0, temp=0.0, seed=0)
    say("large perfect lattice n=", sd.n)
    eng = engine.Engine(sd, ivol=1)
    eng.linkup()
    eng.md_init()
    pot = oracle.Potential()
    r1, r2 = a * np.sqrt(3) / 2, a
    V1, V2 = pot.f("vee", [r1, r2])
    P1, P2 = pot.f("phi", [r1, r2])
    e_atom = 0.5 * (8 * V1 + 6 * V2) - np.sqrt(8 * P1 + 6 * P2)
    assert abs(eng.scalar("pe") / sd.n - e_atom) < 1e-11
    assert np.max(np.abs(eng.forces_cart())) < 1e-10
    eng.close()


def test_gpu_large_random_properties():
    """524,288 atoms, displaced: Newton's third law and translation invariance."""
    sd = lattice.make_system(nc=64, a=2.8665, sigma=0.05, temp=300.0, seed=8)
    say("large random n=", sd.n)
    eng = engine.Engine(sd, ivol=0, press_gpa=0.0)
    eng.linkup()
    eng.md_init()
    f = np.stack([eng.array(n) for n in ("fx", "fy", "fz")], axis=1)
    pe = eng.scalar("pe")
    assert np.all(np.abs(f.sum(axis=0)) < 1e-9 * np.abs(f).max() * np.sqrt(sd.n))
    eng.close()
    shifted = lattice.System(b0=sd.b0, s=sd.s + np.array([0.1234, -0.377, 0.05]), x1=sd.x1,
                             mass=sd.mass, z=sd.z, deltat_fs=sd.deltat_fs)
    eng2 = engine.Engine(shifted, ivol=0)
    eng2.linkup()
    eng2.md_init()
    assert abs(eng2.scalar("pe") - pe) < 1e-11 * abs(pe)
    f2 = np.stack([eng2.array(n) for n in ("fx", "fy", "fz")], axis=1)
    assert rel(f2, f) < 1e-10
    eng2.close()


def test_gpu_fp32_interval_table(oracle):
    """The FP32 force pass's V, V' and phi' (one polynomial per interval
    between the pair knots, the density knots and r*, found through a uniform
    bin grid) against the exact FP64 functions: within 1.5e-6 (V), 4.5e-6 (V')
    and 1.5e-6 eV/A (phi') absolute over 1.91-4.1 A, where |V'| reaches 7.3
    eV/A; exactly zero beyond the outermost knots, and phi' exactly zero below
    the density maximum r*."""
    F = load_golden("c0_nve_2000")
    say("fp32 interval table")
    eng = make_engine(F, precision=engine.PREC_FP32)
    pot = oracle.Potential()
    r = np.linspace(1.91, 4.1, 20001)
    for nm, ref, tol in (("vee32", "vee", 1.5e-6), ("dvee32", "dvee", 4.5e-6), ("dphi32", "dphi", 1.5e-6)):
        got = eng.eval_potential(nm, r)
        err = np.max(np.abs(got - pot.f(ref, r)))
        say(f"{nm}: max abs error {err:.3e}")
        assert err <= tol, (nm, err)
    far = np.array([3.3825, 3.5, 3.72646, 4.0])
    assert not eng.eval_potential("vee32", far).any() and not eng.eval_potential("dvee32", far).any()
    assert not eng.eval_potential("dphi32", np.array([3.72646, 3.9, 5.0])).any()
    rstar = pot.phi_cut[0]
    assert not eng.eval_potential("dphi32", np.array([1.7, 1.8, rstar - 1e-4])).any()
    assert eng.eval_potential("dphi32", np.array([rstar + 1e-3]))[0] != 0.0
    eng.close()


def test_gpu_fp32_close_knots_fall_back_to_knot_sums(oracle):
    """A 6+2 knot set whose first two pair knots lie 0.003 A apart, closer than
    the finest bin of the FP32 interval table: the table is refused (its
    functions are not offered), the FP32 force pass sums the knots directly,
    and its forces stay within 1e-5 eV/A of the oracle's (FP64).  Its energy
    error is larger than the table's (6.3e-6 eV/atom measured: the FP32 knot
    sums cancel terms up to ten times V), bound here at 1e-5 eV/atom."""
    from moldy_amd import potential
    c = potential.coeffs_from_theta(potential.theta_vector())
    c["rv"][1] = c["rv"][0] - 0.003
    sd = lattice.make_system(nc=8, a=2.8665, sigma=0.08, temp=300.0, deltat_fs=1.0, seed=5)
    say("fp32 form, close knots")
    eng = engine.Engine(sd, ivol=1, precision=engine.PREC_FP32, coeffs=c)
    with pytest.raises(engine.MdkError):
        eng.eval_potential("vee32", np.array([2.5]))
    eng.linkup()
    eng.md_init()
    S = oracle.OracleSystem(sd, oracle.Potential(c), ivol=1)
    S.call("linkup")
    S.call("md_init")
    err = float(np.max(np.abs(eng.forces_cart() - S.forces_cart())))
    say(f"close knots: fp32 max force error {err:.3e} eV/A")
    assert err <= 1e-5
    assert abs(eng.scalar("pe") - S.scalar("pe")) / S.n < 1e-5
    eng.close()
