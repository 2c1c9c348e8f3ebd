"""Fermion action with a hand-written adjoint: torch.autograd.Function.

S_f(phi) = Re psi^dag (M^dag M)^{-1} psi per chain; dS_f/dphi = -Ft where Ft is
the adjoint force kernel's output (the force of force_f_fast / force_f_sparse,
hmc_sampler_batch.py:2150-2330).  psi is treated as a constant (as in HMC,
where it is fixed along a trajectory).  With recompute=True the solution xi is
not kept between forward and backward: backward solves again (the memory-lean
choice for the largest lattices); otherwise xi is saved.

phi: [nchain, Ltau*L*L*2] float64 (the kernels' layout); psi: [nchain, Ltau*L*L]
complex128.  precision "fp64" or "mixed" selects the solver.  A solve whose
status is non-finite or a breakdown (kernels.check_status) raises QfError
instead of returning an action and gradient built on a failed solution.
"""
import torch

from . import kernels


def _solve(phi, psi, L, Ltau, dtau, ptab, rtol, max_iter, precision, ru_period):
    status = torch.empty(psi.numel() // (L * L * Ltau), dtype=torch.int32, device=psi.device)
    if precision == "mixed":
        xi, _, err = kernels.pcg_mixed_solve(phi, psi, L, Ltau, dtau, ptab=ptab, rtol=rtol, max_iter=max_iter,
                                             ru_period=ru_period, status=status)
    else:
        xi, _, err = kernels.pcg_solve(phi, psi, L, Ltau, dtau, ptab=ptab, rtol=rtol, max_iter=max_iter, status=status)
    kernels.check_status(status, "FermionAction solve")
    return xi, err


class FermionAction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, psi, L, Ltau, dtau=0.1, ptab=None, rtol=1e-12, max_iter=4000, precision="fp64", recompute=True,
                ru_period=50):
        phi = phi.detach().contiguous()
        psi = psi.detach().contiguous()
        xi, err = _solve(phi, psi, L, Ltau, dtau, ptab, rtol, max_iter, precision, ru_period)
        S = torch.real(torch.sum(psi.conj() * xi, dim=-1))
        ctx.cfg = (L, Ltau, dtau, ptab, rtol, max_iter, precision, recompute, ru_period)
        ctx.err = err
        if recompute:
            ctx.save_for_backward(phi, psi)
        else:
            ctx.save_for_backward(phi, xi)
        return S

    @staticmethod
    def backward(ctx, gS):
        L, Ltau, dtau, ptab, rtol, max_iter, precision, recompute, ru_period = ctx.cfg
        if recompute:
            phi, psi = ctx.saved_tensors
            xi, _ = _solve(phi, psi, L, Ltau, dtau, ptab, rtol, max_iter, precision, ru_period)
        else:
            phi, xi = ctx.saved_tensors
        Ft = kernels.fermion_force(phi, xi, L, Ltau, dtau)
        grad_phi = -Ft * gS.reshape(-1, 1).to(Ft.dtype)
        return grad_phi, None, None, None, None, None, None, None, None, None, None


def fermion_action(phi, psi, L, Ltau, dtau=0.1, ptab=None, rtol=1e-12, max_iter=4000, precision="fp64", recompute=True,
                   ru_period=50):
    """S_f per chain, differentiable in phi through the adjoint force kernel."""
    return FermionAction.apply(phi, psi, L, Ltau, dtau, ptab, rtol, max_iter, precision, recompute, ru_period)
