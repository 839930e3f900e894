"""Adjoint kernel: d(loss)/d(16 knot parameters) over a batch of cells.

loss = sum_c (E_c - E_t,c)^2 / N_c^2 + 0.1 mean_i |F_i - F_t,i|^2 with targets
from the same potential with knots perturbed by a seeded +-5 %.
Checks (FP64 form): against central finite differences (h = 1e-6 relative)
of the GPU loss, to 1e-6 relative; against the FP64 reference gradient from
central differences of the oracle's cell evaluation, to 1e-6 relative; the
vector-Jacobian product through a torch.autograd.Function against torch's
gradcheck.  The FP32-pair form states its gradient error against FP64
(<= 1e-4 of the largest component).  Components are compared relative to
max(|g_k|, 1e-3 max|g|)."""
import numpy as np
import pytest
import torch

from moldy_amd import engine, lattice, potential

pytestmark = pytest.mark.gpu


def _cells():
    # sigma 0.25 A puts pairs into the short-range join (< 1.9 A) and under the flat density cap
    return lattice.make_batch(ncell=12, seed=3, nmin=3, nmax=5, sigma=0.25)


def _targets(cells, seed=11):
    rng = np.random.default_rng(seed)
    th = potential.theta_vector() * (1 + rng.uniform(-0.05, 0.05, 16))
    be = engine.BatchEngine(cells, precision=engine.PREC_FP64, coeffs=potential.coeffs_from_theta(th))
    be.evaluate()
    e, _, f, _ = be.results()
    be.close()
    return e, f


def _rel(g, ref):
    return np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())


def _engine(cells, et, ft, prec):
    be = engine.BatchEngine(cells, precision=prec)
    be.set_targets(et, ft)
    return be


def test_adjoint_matches_finite_differences():
    cells = _cells()
    et, ft = _targets(cells)
    print("[gpu-test] adjoint vs GPU finite differences", sum(c.n for c in cells), flush=True)
    be = _engine(cells, et, ft, engine.PREC_FP64)
    th0 = potential.theta_vector()
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
    assert _rel(g, fd).max() < 1e-6, (g, fd)


def test_adjoint_matches_oracle_reference_gradient(oracle):
    cells = _cells()
    et, ft = _targets(cells)
    print("[gpu-test] adjoint vs oracle FP64 gradient", flush=True)
    be = _engine(cells, et, ft, engine.PREC_FP64)
    loss_gpu, g = be.loss_grad()
    be.close()
    natom = sum(c.n for c in cells)
    off = np.cumsum([0] + [c.n for c in cells])

    def oracle_loss(th):
        pot = oracle.Potential(potential.coeffs_from_theta(th))
        rc = max(th[6:12].max(), th[14:16].max()) + 1e-9
        s = 0.0
        for k, c in enumerate(cells):
            pe, f, _, _ = oracle.eval_cell(pot, c, rc)
            s += (pe - et[k]) ** 2 / c.n ** 2 + 0.1 / natom * ((f - ft[off[k]:off[k + 1]]) ** 2).sum()
        return s

    th0 = potential.theta_vector()
    assert abs(oracle_loss(th0) - loss_gpu) < 1e-10 * loss_gpu
    ref = np.zeros(16)
    for k in range(16):
        h = 1e-6 * abs(th0[k])
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        ref[k] = (oracle_loss(tp) - oracle_loss(tm)) / (2 * h)
    assert _rel(g, ref).max() < 1e-6, (g, ref)


def test_vjp_through_autograd_function():
    from torch_wrap import batch_loss
    cells = _cells()
    et, ft = _targets(cells)
    print("[gpu-test] vjp via torch.autograd.Function", flush=True)
    be = _engine(cells, et, ft, engine.PREC_FP64)
    _, g_fused = be.loss_grad()
    theta = torch.tensor(potential.theta_vector(), dtype=torch.float64, requires_grad=True)
    e_t, f_t = torch.from_numpy(et), torch.from_numpy(ft)
    loss = batch_loss(theta, be, e_t, f_t)
    loss.backward()
    np.testing.assert_allclose(theta.grad.numpy(), g_fused, rtol=1e-10, atol=1e-12 * np.abs(g_fused).max())
    # torch's own central differences on the scalar loss
    assert torch.autograd.gradcheck(lambda t: batch_loss(t, be, e_t, f_t), (theta.detach().requires_grad_(),),
                                    eps=1e-6, atol=1e-6 * float(np.abs(g_fused).max()), rtol=1e-5)
    be.close()


def test_fp32_form_gradient_error():
    cells = _cells()
    et, ft = _targets(cells)
    print("[gpu-test] adjoint FP32 form", flush=True)
    g = {}
    for prec in (engine.PREC_FP64, engine.PREC_FP32):
        be = _engine(cells, et, ft, prec)
        g[prec] = be.loss_grad()[1]
        be.close()
    err = np.abs(g[engine.PREC_FP32] - g[engine.PREC_FP64]).max() / np.abs(g[engine.PREC_FP64]).max()
    print("[gpu-test] FP32 gradient error (max abs / max |g|):", err, flush=True)
    assert err < 1e-4


def test_vjp_needs_forward():
    cells = _cells()[:3]
    be = engine.BatchEngine(cells, precision=engine.PREC_FP64)
    with pytest.raises(engine.MdkError):
        be.vjp(np.zeros(be.ncell), np.zeros((be.natom, 3)))
    be.close()
