"""CgHmcSampler: the PCG convergence harness of qed_fermion/cg_hmc_sampler.py (25-254).

benchmark(bs) builds the reference's field families -- uniform fields of
growing amplitude, the staggered pi flux, and pi flux plus Gaussian noise of
growing width (cg_hmc_sampler.py:106-123) -- and runs the PCG with and without
the preconditioner on them (218-219: rtol=1e-7, max_iter=200).  Here all
fields go through one batched launch per variant instead of a Python loop over
torch.sparse matrices.  Returns the reference's tuple (condition-number
entries stay None as in the reference, whose estimates are commented out) and
keeps the per-field numbers in ``self.last_benchmark``.
"""
import math

import torch

from . import kernels
from .hmc_sampler import HmcSampler


class CgHmcSampler(HmcSampler):
    def __init__(self, J=0.5, Nstep=3000, config=None, **kw):
        kw.setdefault("cuda_graph", False)
        super().__init__(J=J, Nstep=Nstep, config=config, **kw)
        self.plt_cg = False
        self.last_benchmark = None

    def _fields(self, bs):
        g = torch.Generator(device="cpu").manual_seed(25)
        shp = (bs, 2, self.Lx, self.Ly, self.Ltau)
        amp = torch.linspace(0.5 * 3.14, 2 * 3.14, bs, dtype=torch.float64).reshape(-1, 1, 1, 1, 1)
        uni = (torch.rand(shp, generator=g, dtype=torch.float64) - 0.5) * amp
        x = torch.arange(self.Lx).view(-1, 1)
        y = torch.arange(self.Ly).view(1, -1)
        sign = (1 - 2 * ((x + y) % 2)).to(torch.float64)
        pi = torch.zeros(1, 2, self.Lx, self.Ly, self.Ltau, dtype=torch.float64)
        pi[:, 0] = (sign * math.pi / 4).unsqueeze(-1)
        pi[:, 1] = (-sign * math.pi / 4).unsqueeze(-1)
        sig = torch.linspace(0.1 * 3.14 / 2, 1 * 3.14 / 2, bs - 1, dtype=torch.float64).reshape(-1, 1, 1, 1, 1)
        noisy = pi + torch.randn((bs - 1,) + shp[1:], generator=g, dtype=torch.float64) * sig
        return torch.cat([uni, pi, noisy]).to(self.device)

    def preconditioned_cg(self, MhM, b, MhM_inv=None, matL=None, rtol=1e-8, max_iter=100, b_idx=None, axs=None,
                          cg_dtype=torch.complex64, boson=None):
        """PCG on M'M (cg_hmc_sampler / hmc_sampler_batch.py:848-930) for the field ``boson``
        ([nb, 2, Lx, Ly, Ltau]); MhM is not materialised -- the operator is the HIP kernel.
        Returns (x, iterations, last error) per field."""
        if boson is None:
            raise ValueError("pass boson=...: the operator is applied matrix-free from the field")
        phi = self._perm(boson).to(torch.float64 if cg_dtype == torch.complex128 else torch.float32)
        ptab = self._ptab() if (MhM_inv is not None or matL is not None) else None
        bb = b.reshape(phi.shape[0], -1).to(cg_dtype).contiguous()
        x, iters, err = kernels.pcg_solve(phi, bb, self.Lx, self.Ltau, self.dtau, ptab=ptab, rtol=rtol, max_iter=max_iter)
        return x, iters.to(torch.int64), err

    def benchmark(self, bs=5):
        fields = self._fields(bs)
        nb = fields.shape[0]
        torch.manual_seed(25)
        b = (torch.randn(1, self.Vs * self.Ltau, dtype=torch.complex64, device=self.device) / math.sqrt(2)).expand(nb, -1)
        if self.precon_table is None:
            self.reset_precon()
        _, it_p, err_p = self.preconditioned_cg(None, b, MhM_inv=self.precon_table, rtol=1e-7, max_iter=200, boson=fields)
        _, it_n, err_n = self.preconditioned_cg(None, b, rtol=1e-7, max_iter=200, boson=fields)
        self.last_benchmark = {"iters_precon": it_p.cpu().tolist(), "iters_plain": it_n.cpu().tolist(),
                               "err_precon": err_p.cpu().tolist(), "err_plain": err_n.cpu().tolist()}
        convergence_steps = [float(v) for v in it_n.cpu().tolist()]
        mean_conv_steps = sum(convergence_steps) / len(convergence_steps)
        residual_errors = err_n.cpu().tolist()
        return mean_conv_steps, None, None, [], None, None, [], [], residual_errors
