"""torch.autograd.Function over the batched engine -- used by the tests only,
to check the adjoint kernel (mdk_batch_vjp) against torch's finite-difference
gradcheck.  The product path is the C ABI; nothing here is needed to use it."""
import numpy as np
import torch

from moldy_amd import potential


class BatchEnergyForces(torch.autograd.Function):
    """theta (16 knot parameters, float64 CPU tensor) -> (E per cell, F (natom,3))."""

    @staticmethod
    def forward(ctx, theta, be):
        be.set_potential(potential.coeffs_from_theta(theta.detach().numpy()))
        be.evaluate()
        e, _, f, _ = be.results()
        ctx.be = be
        ctx.theta = theta.detach().numpy().copy()
        return torch.from_numpy(e.copy()), torch.from_numpy(f.copy())

    @staticmethod
    def backward(ctx, ge, gf):
        be = ctx.be
        # the product is taken at this forward's potential (the engine may have
        # evaluated other parameters since, e.g. inside gradcheck)
        be.set_potential(potential.coeffs_from_theta(ctx.theta))
        be.evaluate()
        ge = np.zeros(be.ncell) if ge is None else ge.detach().numpy()
        gf = np.zeros((be.natom, 3)) if gf is None else gf.detach().numpy()
        g = be.vjp(ge, gf)
        return torch.from_numpy(g), None


def batch_loss(theta, be, e_target, f_target):
    """sum_c (E_c - E_t,c)^2 / N_c^2 + 0.1 mean_i |F_i - F_t,i|^2 in torch."""
    e, f = BatchEnergyForces.apply(theta, be)
    n = torch.from_numpy(be.natoms.astype(np.float64))
    return (((e - e_target) / n) ** 2).sum() + 0.1 * ((f - f_target) ** 2).sum() / be.natom
