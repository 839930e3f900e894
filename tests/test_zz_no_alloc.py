"""mdk_batch_loss_grad_noalloc: the loss and knot gradient of a batch of cells
written into caller device memory with no allocation (work-queue counter and
per-cell partials in the library's static device scratch).

Bar: bit-exact against mdk_batch_loss_grad (the same kernels in the same
order), on two calls in a row, the second with a smaller batch, so stale
scratch contents from the first call would show; a batch above the scratch
limit is refused before anything is launched."""
import numpy as np
import pytest
import torch

from moldy_amd import engine, lattice, potential

pytestmark = pytest.mark.gpu


def _targets(cells, seed):
    rng = np.random.default_rng(seed)
    th = potential.theta_vector() * (1 + rng.uniform(-0.05, 0.05, 16))
    be = engine.BatchEngine(cells, precision=engine.PREC_FP64, coeffs=potential.coeffs_from_theta(th))
    be.evaluate()
    e, _, f, _ = be.results()
    be.close()
    return e, f


def _check(cells, seed, prec):
    et, ft = _targets(cells, seed)
    be = engine.BatchEngine(cells, precision=prec)
    be.set_targets(et, ft)
    loss0, g0 = be.loss_grad()
    dev = torch.device("cuda:0")
    loss = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
    grad = torch.full((be.ngrad,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    be.loss_grad_noalloc(loss.data_ptr(), grad.data_ptr())
    be.synchronize()
    l1 = float(loss.cpu().numpy()[0])
    g1 = grad.cpu().numpy()
    be.close()
    print("[gpu-test] no-alloc loss", len(cells), "cells:", l1, loss0, flush=True)
    assert loss0 > 0
    assert l1 == loss0
    np.testing.assert_array_equal(g1, g0)


def test_noalloc_loss_grad_matches_loss_grad_twice():
    print("[gpu-test] no-alloc loss/grad, 16 cells then 5", flush=True)
    _check(lattice.make_batch(ncell=16, seed=5, nmin=3, nmax=6, sigma=0.1), 21, engine.PREC_FP32)
    _check(lattice.make_batch(ncell=5, seed=6, nmin=3, nmax=4, sigma=0.1), 22, engine.PREC_FP32)


def test_noalloc_loss_grad_fp64():
    print("[gpu-test] no-alloc loss/grad, FP64 form", flush=True)
    _check(lattice.make_batch(ncell=7, seed=8, nmin=3, nmax=5, sigma=0.1), 23, engine.PREC_FP64)


def test_noalloc_refuses_batches_above_the_limit():
    limit = 8192  # MDK_NOALLOC_MAX_CELLS
    print("[gpu-test] no-alloc limit", limit + 1, "cells", flush=True)
    cells = lattice.make_batch(ncell=limit + 1, seed=9, nmin=3, nmax=3, sigma=0.0)
    be = engine.BatchEngine(cells, precision=engine.PREC_FP32)
    be.set_targets(np.zeros(be.ncell), np.zeros((be.natom, 3)))
    loss = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    grad = torch.zeros(be.ngrad, dtype=torch.float64, device="cuda:0")
    with pytest.raises(engine.MdkError) as ei:
        be.loss_grad_noalloc(loss.data_ptr(), grad.data_ptr())
    be.close()
    assert ei.value.code == 3 and "no-allocation limit" in str(ei.value)
