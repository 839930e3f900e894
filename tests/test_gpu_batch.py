"""Batched engine (one persistent hipGraph-captured launch over many cells)
against the oracle's brute-force single-cell evaluation.

Tolerances: FP64 form -- energies 1e-12 relative, forces and virials 1e-10
relative (norm-wise); FP32-pair/FP64-accumulate form -- max force error
<= 1e-5 eV/A, energy error <= 1e-6 eV/atom, stress error <= 1e-4 GPa."""
import numpy as np
import pytest

from moldy_amd import engine, lattice

pytestmark = pytest.mark.gpu
EV_A3_TO_GPA = 160.21766208


def small_batch(seed=3, ncell=24):
    return lattice.make_batch(ncell=ncell, seed=seed, nmin=3, nmax=6)


def oracle_results(oracle, cells):
    pot = oracle.Potential()
    out = []
    for c in cells:
        pe, f, rho, w = oracle.eval_cell(pot, c)
        out.append((pe, f, rho, w))
    return out


def unpack(res, cells):
    e, w, f, rho = res
    off = np.cumsum([0] + [c.n for c in cells])
    return e, w, [f[off[k]:off[k + 1]] for k in range(len(cells))], [rho[off[k]:off[k + 1]] for k in range(len(cells))]


@pytest.mark.parametrize("prec", [engine.PREC_FP64, engine.PREC_FP32])
def test_batch_matches_oracle(oracle, prec):
    cells = small_batch()
    print("[gpu-test] batch vs oracle prec", prec, "atoms", sum(c.n for c in cells), flush=True)
    be = engine.BatchEngine(cells, precision=prec)
    be.evaluate()
    e, w, fs, rhos = unpack(be.results(), cells)
    ref = oracle_results(oracle, cells)
    for k, c in enumerate(cells):
        pe, f, rho, w9 = ref[k]
        W = np.array([[w[k, 0], w[k, 3], w[k, 4]], [w[k, 3], w[k, 1], w[k, 5]], [w[k, 4], w[k, 5], w[k, 2]]])
        if prec == engine.PREC_FP64:
            assert abs(e[k] - pe) < 1e-12 * abs(pe), (k, e[k], pe)
            assert np.max(np.abs(fs[k] - f)) < 1e-10 * np.max(np.abs(f)), k
            assert np.max(np.abs(W - w9)) < 1e-10 * np.max(np.abs(w9)), k
            np.testing.assert_allclose(rhos[k], rho, rtol=1e-12)
        else:
            assert np.max(np.abs(fs[k] - f)) <= 1e-5, (k, np.max(np.abs(fs[k] - f)))
            assert abs(e[k] - pe) / c.n < 1e-6, k
            vol = abs(np.linalg.det(c.b0))
            assert np.max(np.abs(W - w9)) / vol * EV_A3_TO_GPA < 1e-4, k
    be.close()


def test_batch_graph_replay_matches_eager():
    cells = small_batch(seed=5, ncell=40)
    print("[gpu-test] batch graph replays", flush=True)
    be = engine.BatchEngine(cells, precision=engine.PREC_FP32)
    be.config(block=256, wg_per_cu=4, graph=False)
    be.evaluate()
    e0, w0, f0, _ = be.results()
    be.config(block=256, wg_per_cu=4, graph=True)
    for rep in range(4):  # replays with eager host work and syncs in between
        be.evaluate()
        e, w, f, _ = be.results()
        np.testing.assert_array_equal(e, e0)
        np.testing.assert_array_equal(w, w0)
        np.testing.assert_array_equal(f, f0)
    be.close()


def test_batch_tuner_keeps_results():
    cells = small_batch(seed=9, ncell=64)
    print("[gpu-test] batch tuner", flush=True)
    be = engine.BatchEngine(cells, precision=engine.PREC_FP32)
    be.evaluate()
    e0, w0, f0, _ = be.results()
    block, wg, ms = be.tune(reps=2)
    assert block in (128, 256, 512) and 1 <= wg <= 8 and ms > 0
    be.evaluate()
    e, w, f, _ = be.results()
    assert np.max(np.abs(e - e0)) < 1e-10 * np.max(np.abs(e0))
    assert np.max(np.abs(f - f0)) < 1e-9
    be.close()
