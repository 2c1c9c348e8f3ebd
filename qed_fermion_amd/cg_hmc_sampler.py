"""CgHmcSampler: the PCG convergence harness of qed_fermion/cg_hmc_sampler.py (25-254).

benchmark(bs) builds the reference's field families in its order -- bs uniform
fields of growing amplitude, the staggered pi flux, and bs-1 pi-flux fields plus
Gaussian noise of growing width (cg_hmc_sampler.py:106-123), 2*bs fields in all --
and, like the reference, studies only the fields with index 4 and 6, in reversed
order (171-172).  Each is solved with and without the preconditioner at rtol=1e-7,
max_iter=200 (218-219).  The operator is the HIP M^dag M applied from the field; the
reference builds a torch.sparse M'M per field instead.

Returns the reference's 9-tuple with its slot meanings (230-234):
  0 mean_conv_steps      None: the reference never fills convergence_steps
  1 mean_sparsity        mean of get_M_sparse's blk_sparsity over the studied fields,
                         which is -1.0 (hmc_sampler_batch.py:1877 sets it for every tau)
  2 mean_condition_num   None (its estimates are commented out)
  3 condition_numbers    []
  4 mean_sig_max         None
  5 mean_sig_min         None
  6 sig_max_values       []
  7 sig_min_values       []
  8 residual_errors      last relative residual of the unpreconditioned CG per studied field
The per-field numbers of both solves are kept in ``self.last_benchmark``.
"""
import math

import torch

from . import kernels
from .hmc_sampler import HmcSampler

# The studied field indices of the reference's benchmark loop (cg_hmc_sampler.py:171-172).
BENCHMARK_FIELDS = (4, 6)


class CgHmcSampler(HmcSampler):
    def __init__(self, J=0.5, Nstep=3000, config=None, **kw):
        kw.setdefault("cuda_graph", False)
        super().__init__(J=J, Nstep=Nstep, config=config, **kw)
        self.plt_cg = False
        self.plt_pattern = False
        # cg_hmc_sampler.py:22 (module-level cg_dtype = torch.complex64)
        self.cg_dtype = torch.complex64
        self.last_benchmark = None

    def benchmark_fields(self, bs):
        """The 2*bs fields of benchmark(bs) (cg_hmc_sampler.py:106-123), [2*bs, 2, Lx, Ly, Ltau] float64."""
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
        """PCG on M'M (cg_hmc_sampler / hmc_sampler_batch.py:848-930) for the fields ``boson``
        ([nb, 2, Lx, Ly, Ltau]); MhM is not materialised -- the operator is the HIP kernel.
        Returns (x, iterations, last error) per field; iterations counts the iterations
        after which the residual was still above rtol (the reference's cnt)."""
        if boson is None:
            raise ValueError("pass boson=...: the operator is applied matrix-free from the field")
        phi = self._perm(boson).to(torch.float64 if cg_dtype == torch.complex128 else torch.float32)
        ptab = self._ptab() if (MhM_inv is not None or matL is not None) else None
        bb = b.reshape(-1, phi.shape[1] // 2).expand(phi.shape[0], -1).to(cg_dtype).contiguous()
        x, iters, err = kernels.pcg_solve(phi, bb, self.Lx, self.Ltau, self.dtau, ptab=ptab, rtol=rtol, max_iter=max_iter)
        return x, iters.to(torch.int64), err

    def benchmark(self, bs=5):
        fields = self.benchmark_fields(bs)
        torch.manual_seed(25)
        b = (torch.randn(self.Vs * self.Ltau, dtype=torch.complex64, device=self.device) / math.sqrt(2))
        if self.precon_table is None:
            self.reset_precon()
        idx = [i for i in reversed(range(2 * bs)) if i in BENCHMARK_FIELDS]
        rec = {"fields": idx, "b": b, "iters_precon": [], "iters_plain": [], "err_precon": [], "err_plain": []}
        blk_sparsities, residual_errors = [], []
        if idx:
            sel = fields[idx]
            _, it_p, err_p = self.preconditioned_cg(None, b, MhM_inv=self.precon_table, rtol=1e-7, max_iter=200,
                                                    cg_dtype=self.cg_dtype, boson=sel)
            _, it_n, err_n = self.preconditioned_cg(None, b, MhM_inv=None, rtol=1e-7, max_iter=200,
                                                    cg_dtype=self.cg_dtype, boson=sel)
            rec.update({"iters_precon": it_p.cpu().tolist(), "iters_plain": it_n.cpu().tolist(),
                        "err_precon": err_p.cpu().tolist(), "err_plain": err_n.cpu().tolist()})
            residual_errors = rec["err_plain"]
            blk_sparsities = [-1.0] * len(idx)   # get_M_sparse's blk_sparsity (hmc_sampler_batch.py:1877)
        self.last_benchmark = rec
        convergence_steps, condition_numbers, sig_max_values, sig_min_values = [], [], [], []
        mean_conv_steps = sum(convergence_steps) / len(convergence_steps) if convergence_steps else None
        mean_condition_num = sum(condition_numbers) / len(condition_numbers) if condition_numbers else None
        mean_sparsity = sum(blk_sparsities) / len(blk_sparsities) if blk_sparsities else None
        mean_sig_max = sum(sig_max_values) / len(sig_max_values) if sig_max_values else None
        mean_sig_min = sum(sig_min_values) / len(sig_min_values) if sig_min_values else None
        return (mean_conv_steps, mean_sparsity, mean_condition_num, condition_numbers, mean_sig_max, mean_sig_min,
                sig_max_values, sig_min_values, residual_errors)
