"""Parity at the full workload sizes.

* 524,288-atom bcc Fe NPT (64^3 cells, sigma 0.05 A, 300 K): the FP32 form's
  forces within 1e-5 eV/A of the oracle (FP64), energy within 1e-6 eV/atom and
  the virial within 1e-4 GPa, at step 0 and again after 100 NPT steps (with
  list rebuilds) on the evolved configuration; the FP64 form within 1e-10
  (forces, norm-wise relative) and 1e-12 (energy) at step 0.
* 3,456,000-atom hot liquid (120^3 cells at 7.0 g/cm^3, 2500 K, dt 0.5 fs,
  skin 0.3 A): the FP32 form's forces within 1e-5 eV/A of the oracle at step 0
  and after 100 steps.
* The 4096-cell batch (seed 7): FP64 and FP32 batch results of one bcc and one
  fcc cell with vacancies for each n = 7..10, from the evaluation of the whole
  batch, against the oracle's single-cell evaluation (FP64: energy 1e-12,
  forces and virial 1e-10 relative, rho 1e-12; FP32: forces 1e-5 eV/A,
  energy 1e-6 eV/atom, stress 1e-4 GPa).
* The adjoint gradient over the whole 4096-cell batch (FP64) against central
  finite differences of the GPU loss (h = 1e-6 relative), 1e-6 relative.
"""
import sys

import numpy as np
import pytest

from moldy_amd import engine, lattice, potential

pytestmark = pytest.mark.gpu
EV_A3_TO_GPA = 160.21766208


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _state(eng, sd):
    """The engine's current configuration as a System (positions folded into [0,1))."""
    s = np.stack([eng.array(n) for n in ("x0", "y0", "z0")], axis=1)
    return lattice.System(b0=eng.tensor("b0"), s=s - np.floor(s), x1=np.zeros_like(s), mass=sd.mass,
                          z=sd.z, deltat_fs=sd.deltat_fs)


def _oracle_first_force(oracle, sd, **kw):
    S = oracle.OracleSystem(sd, oracle.Potential(), **kw)
    S.call("linkup")
    S.call("md_init")
    return S


def _fp32_bounds(eng, S, what):
    err = np.max(np.abs(eng.forces_cart() - S.forces_cart()))
    de = abs(eng.scalar("pe") - S.scalar("pe")) / S.n
    dsig = np.max(np.abs(eng.tensor("tp") - S.tensor("tp"))) / S.scalar("vol") * EV_A3_TO_GPA
    say(f"{what}: fp32 max force error {err:.3e} eV/A, energy {de:.3e} eV/atom, virial {dsig:.3e} GPa")
    assert err <= 1e-5
    assert de < 1e-6
    assert dsig < 1e-4


def bcc_npt_524288():
    return lattice.make_system(nc=64, a=2.8665, sigma=0.05, temp=300.0, deltat_fs=1.0, seed=2026)


def test_fullsize_npt_524288_step0(oracle):
    sd = bcc_npt_524288()
    say("524,288-atom NPT, step 0: oracle")
    S = _oracle_first_force(oracle, sd, ivol=0, ntcm=10)
    for prec in (engine.PREC_FP32, engine.PREC_FP64):
        say("524,288-atom NPT, step 0: precision", prec)
        eng = engine.Engine(sd, ivol=0, press_gpa=0.0, boxmass=50.0, ntcm=10, precision=prec, rebuild=1)
        eng.linkup()
        eng.md_init()
        if prec == engine.PREC_FP32:
            _fp32_bounds(eng, S, "npt step 0")
        else:
            f, fr = eng.forces_cart(), S.forces_cart()
            assert np.max(np.abs(f - fr)) / np.max(np.abs(fr)) < 1e-10
            assert abs(eng.scalar("pe") - S.scalar("pe")) < 1e-12 * abs(S.scalar("pe"))
        eng.close()


def test_fullsize_npt_524288_after_100_steps(oracle):
    sd = bcc_npt_524288()
    say("524,288-atom NPT: 100 FP32 steps")
    eng = engine.Engine(sd, ivol=0, press_gpa=0.0, boxmass=50.0, ntcm=10, precision=engine.PREC_FP32, rebuild=1)
    eng.linkup()
    eng.md_init()
    eng.md_steps(100, trace=False)
    rebuilds, _ = eng.stats()
    say("rebuilds", rebuilds, "; oracle on the evolved configuration")
    S = _oracle_first_force(oracle, _state(eng, sd), ivol=0, ntcm=10)
    assert rebuilds >= 2
    _fp32_bounds(eng, S, "npt step 100")
    eng.close()


def liquid_3456000():
    a = lattice.liquid_lattice_parameter(7.0)
    return lattice.make_system(nc=120, a=a, sigma=0.05, temp=2500.0, deltat_fs=0.5, seed=2500)


@pytest.mark.parametrize("steps", [0, 100])
def test_fullsize_hot_liquid_3456000(oracle, steps):
    sd = liquid_3456000()
    say(f"3,456,000-atom hot liquid, FP32, {steps} steps")
    eng = engine.Engine(sd, ivol=1, ntcm=1000, precision=engine.PREC_FP32, rebuild=1, skin=0.3)
    eng.linkup()
    eng.md_init()
    if steps:
        eng.md_steps(steps, trace=False)
        say("rebuilds", eng.stats()[0])
    say("oracle")
    S = _oracle_first_force(oracle, _state(eng, sd) if steps else sd, ivol=1, ntcm=1000)
    _fp32_bounds(eng, S, f"liquid step {steps}")
    eng.close()


def _sampled(cells):
    """One bcc and one fcc cell with vacancies for each n = 7..10."""
    out = []
    for nc in range(7, 11):
        for fcc in (False, True):
            out.append(next(k for k, c in enumerate(cells)
                            if c.info["nc"] == nc and c.info["fcc"] == fcc and c.info["vacancy"] > 0.005))
    return out


@pytest.mark.parametrize("prec", [engine.PREC_FP64, engine.PREC_FP32])
def test_fullsize_batch_4096_sampled_cells(oracle, prec):
    cells = lattice.make_batch(ncell=4096, seed=7)
    sel = _sampled(cells)
    say("4096-cell batch, precision", prec, "cells", sel)
    be = engine.BatchEngine(cells, precision=prec)
    be.evaluate()
    e, w, f, rho = be.results()
    be.close()
    off = np.cumsum([0] + [c.n for c in cells])
    pot = oracle.Potential()
    for k in sel:
        c = cells[k]
        pe, fr, rr, w9 = oracle.eval_cell(pot, c)
        fk, rk = f[off[k]:off[k + 1]], rho[off[k]:off[k + 1]]
        W = np.array([[w[k, 0], w[k, 3], w[k, 4]], [w[k, 3], w[k, 1], w[k, 5]], [w[k, 4], w[k, 5], w[k, 2]]])
        if prec == engine.PREC_FP64:
            assert abs(e[k] - pe) < 1e-12 * abs(pe), (k, e[k], pe)
            assert np.max(np.abs(fk - fr)) < 1e-10 * np.max(np.abs(fr)), k
            assert np.max(np.abs(W - w9)) < 1e-10 * np.max(np.abs(w9)), k
            np.testing.assert_allclose(rk, rr, rtol=1e-12)
        else:
            err = np.max(np.abs(fk - fr))
            say(f"cell {k} (n={c.info['nc']}, {'fcc' if c.info['fcc'] else 'bcc'}, {c.n} atoms): "
                f"fp32 force error {err:.3e}")
            assert err <= 1e-5, (k, err)
            assert abs(e[k] - pe) / c.n < 1e-6, k
            assert np.max(np.abs(W - w9)) / abs(np.linalg.det(c.b0)) * EV_A3_TO_GPA < 1e-4, k


def test_fullsize_batch_4096_adjoint_finite_differences():
    cells = lattice.make_batch(ncell=4096, seed=7)
    say("4096-cell adjoint vs central differences")
    rng = np.random.default_rng(11)
    th_t = potential.theta_vector() * (1 + rng.uniform(-0.05, 0.05, 16))
    be = engine.BatchEngine(cells, precision=engine.PREC_FP64, coeffs=potential.coeffs_from_theta(th_t))
    be.evaluate()
    et, _, ft, _ = be.results()
    th0 = potential.theta_vector()
    be.set_potential(potential.coeffs_from_theta(th0))
    be.set_targets(et, ft)
    loss0, g = be.loss_grad()
    assert loss0 > 0
    fd = np.zeros(16)
    for k in range(16):
        h = 1e-6 * abs(th0[k])
        vals = []
        for sgn in (1, -1):
            t = th0.copy()
            t[k] += sgn * h
            be.set_potential(potential.coeffs_from_theta(t))
            vals.append(be.loss_grad()[0])
        fd[k] = (vals[0] - vals[1]) / (2 * h)
    be.close()
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    say(f"max relative difference {rel.max():.3e}")
    assert rel.max() < 1e-6, (g, fd)
