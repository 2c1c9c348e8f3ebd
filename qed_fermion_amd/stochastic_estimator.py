"""StochaticEstimator: the multi-RHS solve path of qed_fermion/stochastic_estimator.py.

Same class and method names (the reference's spelling), same tensor shapes:

  random_vec_bin()          -> eta [Nrv, Ltau*Ly*Lx], entries (+-1 +- i)/sqrt(2)      (210-224)
  set_eta_G_eta(boson, eta) -> self.G_eta = (M^dag M)^{-1} M^dag eta = M^{-1} eta      (409-433)

  T1 / T21 / T2 / T4 / Tvv(eta, G_eta, boson)   bond-bond estimators    (1393-1654)
  get_bond_bond_per_b2(boson), get_fermion_obsr_compile(bosons, eta)    (2492-2641)
  bond_corr, reorder_fft_grid2[_inverse]                                (253-262, 2643-2651)

MI355X differences: the reference repeats the boson field Nrv times
(``boson.repeat(Nrv, 1)``, 429) so its batched kernels see one field per RHS;
here all Nrv right-hand sides share ONE field (phi stride 0), so M^dag and
every PCG iteration read phi / its hop table once per launch for the whole
batch, in one multi-RHS launch per iteration.  The bond-bond estimators run
on HIP kernels around rocFFT transforms (qed_fermion_amd/observables.py):
chunks of pairs instead of batches of int(0.1 Nrv), k-space products summed
over all pairs and slices in FP64, one inverse transform per estimator.
"""
import math
import os

import torch

from . import kernels, observables
from .fermion_obsr_graph_runner import (GEtaGraphRunner, T1GraphRunner, T2GraphRunner, T4GraphRunner, T21GraphRunner,
                                        TvvGraphRunner)


class StochaticEstimator:
    def __init__(self, hmc, cuda_graph_se=False, Nrv=None, max_iter_se=None):
        self.hmc_sampler = hmc
        self.Nrv = int(Nrv if Nrv is not None else os.getenv("Nrv", "20"))
        self.max_iter_se = int(max_iter_se if max_iter_se is not None else os.getenv("max_iter_se", "100"))
        self.Lx, self.Ly, self.Ltau = hmc.Lx, hmc.Ly, hmc.Ltau
        self.dtau = hmc.dtau
        self.Vs = hmc.Vs
        self.cuda_graph_se = cuda_graph_se
        self.device = hmc.device
        self.dtype = hmc.dtype
        self.cdtype = hmc.cdtype
        self.graph_memory_pool = hmc.graph_memory_pool
        self.G_eta_graph_runner = GEtaGraphRunner(self)
        self.T1_graph_runner = T1GraphRunner(self)
        self.T21_graph_runner = T21GraphRunner(self)
        self.T2_graph_runner = T2GraphRunner(self)
        self.T4_graph_runner = T4GraphRunner(self)
        self.Tvv_graph_runner = TvvGraphRunner(self)
        self._bb_est = None
        self.num_samples = lambda nrv: math.comb(nrv, 2)
        self.batch_size = lambda nrv: int(nrv * 0.1)
        self.G_eta = None
        self.eta = None
        if hmc.precon_table is None and hmc.dtau <= 0.1 and hmc.precon_on:
            hmc.reset_precon()

    def initialize(self):
        self.indices_r2 = torch.combinations(torch.arange(self.Nrv, device=self.device), r=2, with_replacement=False)

    def init_cuda_graph(self):
        """Capture the G_eta solve and the five estimators (stochastic_estimator.py:111-200); the
        solve is captured with the sampler's first CG gear as its budget (134-141)."""
        hmc = self.hmc_sampler
        if self.cuda_graph_se:
            self.graph_memory_pool = self.G_eta_graph_runner.capture(
                max_iter_se=hmc._MAX_ITERS_TO_CAPTURE[0],
                graph_memory_pool=self.graph_memory_pool)
            for r in (self.T1_graph_runner, self.T21_graph_runner, self.T2_graph_runner, self.T4_graph_runner,
                      self.Tvv_graph_runner):
                self.graph_memory_pool = r.capture(self.graph_memory_pool)

    def random_vec_bin(self):
        """Nrv Z2 x Z2 noise vectors (+-1 +- i)/sqrt(2), [Nrv, Lx*Ly*Ltau] (stochastic_estimator.py:210-224)."""
        n = self.Lx * self.Ly * self.Ltau
        re = torch.randint(0, 2, (self.Nrv, n), device=self.device).to(torch.float64) * 2 - 1
        im = torch.randint(0, 2, (self.Nrv, n), device=self.device).to(torch.float64) * 2 - 1
        return torch.complex(re, im).div_(math.sqrt(2)).to(self.cdtype)

    def set_eta_G_eta(self, boson, eta):
        """G_eta = M^{-1} eta for all Nrv noise vectors (stochastic_estimator.py:409-425)."""
        self.eta = eta
        if self.cuda_graph_se and self.G_eta_graph_runner.graph is not None:
            G_eta = self.G_eta_graph_runner(boson, eta)
        else:
            G_eta = self.set_eta_G_eta_inner(boson, eta)
        self.G_eta = G_eta

    def set_eta_G_eta_inner(self, boson, eta):
        """psi = M^dag eta; G_eta = (M^dag M)^{-1} psi with one shared field (stochastic_estimator.py:427-433).
        The solve runs the sampler's max_iter iterations at most, as Ot_inv_psi_fast does there
        (GEtaGraphRunner.capture sets it to max_iter_se while it captures)."""
        hmc = self.hmc_sampler
        phi = boson.permute([0, 4, 3, 2, 1]).reshape(1, -1)
        real = torch.float32 if eta.dtype == torch.complex64 else torch.float64
        phi = phi.to(real).contiguous()
        eta = eta.contiguous()
        psi = kernels.apply("Mdag", phi, eta, self.Lx, self.Ltau, self.dtau)
        ptab = hmc._ptab()
        if hmc.precision == "mixed":
            G_eta, _, _ = kernels.pcg_mixed_solve(phi.double(), psi.to(torch.complex128), self.Lx, self.Ltau, self.dtau,
                                                  ptab=ptab, rtol=hmc.cg_rtol, max_iter=hmc.max_iter,
                                                  ru_period=hmc.ru_period)
        else:
            G_eta, _, _ = kernels.pcg_solve(phi, psi, self.Lx, self.Ltau, self.dtau, ptab=ptab, rtol=hmc.cg_rtol,
                                            max_iter=hmc.max_iter)
        return G_eta

    # ---------------------------------------------------------- bond-bond estimators
    def _bb(self, eta):
        est = self._bb_est
        if est is None or est.dtype != eta.dtype or est.nrv != eta.shape[0]:
            est = observables.BondBondEstimator(self.Lx, self.Ltau, eta.shape[0], dtype=eta.dtype, device=self.device)
            self._bb_est = est
        return est

    def T1(self, eta, G_eta, boson):
        """stochastic_estimator.py:1393-1445; boson [1, 2, Lx, Ly, Ltau]; returns [Ly, Lx]."""
        return self._bb(eta).T("T1", eta, G_eta, boson)

    def T21(self, eta, G_eta, boson):
        """stochastic_estimator.py:1447-1502."""
        return self._bb(eta).T("T21", eta, G_eta, boson)

    def T2(self, eta, G_eta, boson):
        """stochastic_estimator.py:1504-1559."""
        return self._bb(eta).T("T2", eta, G_eta, boson)

    def T4(self, eta, G_eta, boson):
        """stochastic_estimator.py:1561-1610."""
        return self._bb(eta).T("T4", eta, G_eta, boson)

    def Tvv(self, eta, G_eta, boson):
        """(Tvv, Tv) (stochastic_estimator.py:1612-1654)."""
        return self._bb(eta).Tvv(eta, G_eta, boson)

    def get_bond_bond_per_b2(self, boson):
        """BB0_r, BB_r, B_r, BB_k of one chain from self.eta / self.G_eta (stochastic_estimator.py:2594-2641)."""
        runners = None
        if self.cuda_graph_se and self.T1_graph_runner.graph is not None:
            runners = {"T1": self.T1_graph_runner, "T21": self.T21_graph_runner, "T2": self.T2_graph_runner,
                       "T4": self.T4_graph_runner, "Tvv": self.Tvv_graph_runner}
        return self._bb(self.eta).bond_bond(self.eta, self.G_eta, boson, runners)

    @torch.inference_mode()
    def get_fermion_obsr_compile(self, bosons, eta):
        """Per chain: G_eta solve, then the bond-bond observables, stacked to [bs, Ly, Lx]
        (stochastic_estimator.py:2492-2522)."""
        obsrs = []
        for b in range(bosons.shape[0]):
            boson = bosons[b].unsqueeze(0)
            self.set_eta_G_eta(boson, eta)
            obsrs.append(self.get_bond_bond_per_b2(boson))
        return {k: torch.stack([o[k] for o in obsrs], dim=0) for k in obsrs[0]}

    def bond_corr(self, BB_r_mean, B_r_mean):
        """stochastic_estimator.py:2643-2651."""
        return observables.bond_corr(BB_r_mean, B_r_mean)

    reorder_fft_grid2 = staticmethod(observables.reorder_fft_grid2)
    reorder_fft_grid2_inverse = staticmethod(observables.reorder_fft_grid2_inverse)
