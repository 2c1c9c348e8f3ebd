"""HmcSampler: the HMC fermion hot path of qed_fermion/hmc_sampler_batch.py on MI355X.

Same class name, method names, argument meanings, tensor layouts and return
tuples as the reference's ``HmcSampler`` for the parts of the sampler on the
fermion hot path (operator, preconditioner, PCG, fermion force, boson forces,
leapfrog, Metropolis), so reference-shaped driver code runs unchanged:

  boson        [bs, 2, Lx, Ly, Ltau]            (reference storage layout)
  fermion vec  [bs, Ltau*Ly*Lx]                  (index tau*Vs + y*Lx + x)
  permuted     boson.permute([0, 4, 3, 2, 1])    ([bs, Ltau, Ly, Lx, 2], the kernels' layout)

The work runs in libqf_hip.so (operators, PCG, force, boson sector); there is
no CPU or torch-sparse fallback.  Configuration is read from the reference's
environment variables when the sampler is built (``L``, ``asym``, ``bs``,
``dtau``, ``K``, ``compact``, ``precon``, ``max_iter``, ``cuda_graph``, ``seed``;
hmc_sampler_batch.py:20-80) and can be overridden by keyword.

``precision``:
  "fp32"  complex64 vectors and float32 boson (the reference's dtypes), FP64 dots
  "mixed" complex64 PCG with FP64 reliable updates, complex128 results
  "fp64"  complex128 / float64 throughout
"""
import collections
import math
import os
import warnings

import torch

from . import checkpoint, fourier_accel, kernels, precon
from .trajectory import TrajectoryEngine, adjust_delta_t, couplings


def _env(name, default, cast):
    return cast(os.getenv(name, str(default)))


class HmcSampler(object):
    """hmc_sampler_batch.HmcSampler (hmc_sampler_batch.py:149-319) on the HIP kernels."""

    def __init__(self, Lx=6, Ltau=10, J=0.5, Nstep=3000, config=None, *, bs=None, dtau=None, K=None, compact=None,
                 precon_on=None, max_iter=None, cg_rtol=1e-9, cuda_graph=None, precision="fp32", ru_period=50,
                 device="cuda", seed=None, ckpt_root=None, max_tau_block_idx=None):
        self.Lx = Lx
        self.Ly = Lx
        self.Ltau = Ltau
        self.Vs = Lx * Lx
        self.bs = bs if bs is not None else _env("bs", 2, int)
        self.Nf = 2
        self.dtau = dtau if dtau is not None else _env("dtau", 0.1, float)
        Kin = K if K is not None else _env("K", 1.0, float)
        # Lagrangian * dtau (hmc_sampler_batch.py:169-175)
        self.J = J / self.dtau * self.Nf / 4
        self.K = Kin * self.dtau * self.Nf * 1 / 2
        self.t = 1
        self.compact = bool(compact) if compact is not None else _env("compact", 0, int) != 0
        self.precon_on = bool(precon_on) if precon_on is not None else _env("precon", 1, int) == 1
        self.N_step = Nstep
        self.m = 1
        self.delta_t = 0.06
        self.N_leapfrog = 5
        self.cg_rtol = cg_rtol
        self.max_iter = max_iter if max_iter is not None else _env("max_iter", 400, int)
        self.cuda_graph = bool(cuda_graph) if cuda_graph is not None else _env("cuda_graph", 1, int) != 0
        if precision not in ("fp32", "mixed", "fp64"):
            raise ValueError(f"precision {precision!r}")
        self.precision = precision
        self.ru_period = ru_period
        self.device = torch.device(device)
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        self.cdtype = torch.complex128 if precision in ("fp64", "mixed") else torch.complex64
        self.use_cuda_kernel = True
        seed = seed if seed is not None else _env("seed", 42, int)
        torch.manual_seed(seed)

        self.delta_t_tensor = torch.full((self.bs,), self.delta_t, device=self.device, dtype=self.dtype)
        self.lower_limit = 0.25
        self.upper_limit = 0.5
        # tau-block updates (hmc_sampler_batch.py:158-162): max_tau_block_idx blocks of Ltau // blocks
        # slices, one Metropolis step each, only when Ltau >= 10 Lx
        nblk = max_tau_block_idx if max_tau_block_idx is not None else _env("max_tau_block_idx", 1, int)
        self.max_tau_block_idx = nblk if self.Ltau // self.Lx // 10 >= 1 else 1
        self.tau_block_idx = 0
        self.tau_block_size = self.Ltau // self.max_tau_block_idx
        self.threshold_queue = [collections.deque(maxlen=5 * self.max_tau_block_idx) for _ in range(self.bs)]
        # adaptive mass estimate (hmc_sampler_batch.py:268-285; the environment knobs of 20-80, 126)
        debug = _env("debug", 0, int) != 0
        self.mass_mode = _env("mass_mode", 0, int)
        self._fa = fourier_accel.SigmaHatState(
            self.bs, self.Lx, self.Ly, self.Ltau, self.device, self.dtype,
            sigma_mini_batch_size=50 if debug else 1000, anneal_times=5 if debug else 10,
            lmd=_env("lmd", 0.9, float), sig_min=_env("sig_min", 0.8, float), sig_max=_env("sig_max", 1.2, float),
            lambda_=_env("lambda_", 0.05, float))
        # CG iteration gears (hmc_sampler_batch.py:305-309): one gear unless set otherwise
        self._MAX_ITERS_TO_CAPTURE = [self.max_iter] if self.dtau <= 0.1 and self.precon_on else [400]

        self.precon = None
        self.precon_csr = None
        self.precon_table = None
        self.force_graph_runners = {}
        self.leapfrog_cmp_graph_runners = None
        self.graph_memory_pool = None
        self._engine = None
        self.step = 0
        self.cur_step = 0
        # Folder holding check_points/ and preconditioners/ (the reference keeps them next to its
        # sources); None: start from the pi-flux configuration and build P in memory.
        self.ckpt_root = ckpt_root
        self.last_solve_status = None   # QF_SOLVE_* codes of the last solve(s), on the device
        self.reset()

    # ------------------------------------------------------------------ setup
    def reset(self):
        """reset (hmc_sampler_batch.py:321-328): geometry, run tag, then the ensemble start
        when a checkpoint root is set (pi flux plus noise otherwise)."""
        self.Vs = self.Lx * self.Ly
        self.initialize_geometry()
        self.initialize_specifics()
        if self.ckpt_root is not None:
            self.initialize_boson_ensemble()
        else:
            self.initialize_boson_pi_flux_randn_matfree()

    def initialize_specifics(self):
        """The run tag used in file names (hmc_sampler_batch.py:330-331)."""
        self.specifics = checkpoint.specifics(self)

    def initialize_boson_ensemble(self, root=None):
        """Start from a saved ensemble member under ``root``/check_points/boson_ensemble
        (hmc_sampler_batch.py:971-1016); falls back to the pi-flux start.  Returns the members found."""
        return checkpoint.initialize_boson_ensemble(self, root if root is not None else self.ckpt_root)

    def save_to_file(self, res, data_folder, filename):
        """save_to_file (hmc_sampler_batch.py:3211-3215)."""
        return checkpoint.save_to_file(res, data_folder, filename)

    def initialize_geometry(self):
        """Checkerboard bond groups (hmc_sampler_batch.py:937-969): i/j lists and boson indices."""
        Lx, Ly, dev = self.Lx, self.Ly, self.device
        xs = torch.arange(0, Lx, 2, device=dev).unsqueeze(0)
        ys = torch.arange(0, Ly, 2, device=dev).unsqueeze(1)
        xs2 = torch.arange(1, Lx, 2, device=dev).unsqueeze(0)
        ys2 = torch.arange(1, Ly, 2, device=dev).unsqueeze(1)
        even = torch.cat([(xs + ys * Lx).reshape(-1), (xs2 + ys2 * Lx).reshape(-1)])
        ex = torch.cat([xs.expand(ys.shape[0], -1).reshape(-1), xs2.expand(ys2.shape[0], -1).reshape(-1)])
        ey = torch.cat([ys.expand(-1, xs.shape[1]).reshape(-1), ys2.expand(-1, xs2.shape[1]).reshape(-1)])
        self.i_list_1, self.j_list_1 = even, (ex + 1) % Lx + ey * Lx
        self.i_list_2, self.j_list_2 = even, ex + ((ey + 1) % Ly) * Lx
        self.i_list_3, self.j_list_3 = (ex - 1) % Lx + ey * Lx, even
        self.i_list_4, self.j_list_4 = ex + ((ey - 1) % Ly) * Lx, even
        self.boson_idx_list_1 = self.i_list_1 * 2
        self.boson_idx_list_2 = self.i_list_2 * 2 + 1
        self.boson_idx_list_3 = self.i_list_3 * 2
        self.boson_idx_list_4 = self.i_list_4 * 2 + 1

    def initialize_boson_pi_flux_randn_matfree(self):
        """Staggered pi flux (x-link sign*pi/4, y-link -sign*pi/4, sign = (-1)^(x+y)) for the last
        chain; the other bs-1 chains add tau-independent Gaussian noise with standard deviations
        linspace(0.02, 0.08, bs-1) (hmc_sampler_batch.py:1026-1043)."""
        x = torch.arange(self.Lx, device=self.device).view(-1, 1)
        y = torch.arange(self.Ly, device=self.device).view(1, -1)
        sign = (1 - 2 * ((x + y) % 2)).to(self.dtype)
        stag = torch.zeros(1, 2, self.Lx, self.Ly, self.Ltau, device=self.device, dtype=self.dtype)
        stag[:, 0] = (sign * math.pi / 4).unsqueeze(-1)
        stag[:, 1] = (-sign * math.pi / 4).unsqueeze(-1)
        noise = torch.randn(self.bs - 1, 2, self.Lx, self.Ly, 1, device=self.device, dtype=self.dtype)
        levels = torch.linspace(0.02, 0.08, self.bs - 1, device=self.device, dtype=self.dtype).view(-1, 1, 1, 1, 1)
        self.boson = torch.cat([stag.repeat(self.bs - 1, 1, 1, 1, 1) + noise * levels, stag], dim=0)

    def initialize_boson_test(self):
        """Zero flux plus Gaussian noise of standard deviation linspace(0.1, 0.5, bs) per chain
        (hmc_sampler_batch.py:1018-1024)."""
        levels = torch.linspace(0.1, 0.5, self.bs, device=self.device, dtype=self.dtype).view(-1, 1, 1, 1, 1)
        self.boson = torch.randn(self.bs, 2, self.Lx, self.Ly, self.Ltau, device=self.device, dtype=self.dtype) * levels

    def initialize_boson_uniform(self):
        self.boson = (torch.rand(self.bs, 2, self.Lx, self.Ly, self.Ltau, device=self.device, dtype=torch.float64) * 2 - 1) \
            .mul_(math.pi).to(self.dtype)

    def reset_precon(self):
        """reset_precon (hmc_sampler_batch.py:461-547): the manual preconditioner as precon_csr
        (complex64 CSR, like the reference) and as the band table the kernels apply.

        With a checkpoint root, P comes from ``preconditioners/precon_ckpt_L_..._t_1.pt`` there
        (verified to be the tau-banded P_tau (x) I); a missing file is written from the manual
        preconditioner first, as the reference does through get_precon_man."""
        if self.ckpt_root is None:
            d = precon.get_precon_man(self.Lx, self.Ltau)
            self.precon_csr = precon.coo_to_csr(d, device=self.device)
            self.precon_table = precon.band_table(self.Ltau, dtype=torch.float64, device=self.device)
            return
        path = os.path.join(self.ckpt_root, "preconditioners",
                            checkpoint.precon_file_name(self.Lx, self.Ltau, self.dtau, self.t))
        if not os.path.exists(path):
            checkpoint.save_precon(precon.get_precon_man(self.Lx, self.Ltau), path)
        tab, csr = checkpoint.precon_table_from_file(path, self.Lx, self.Ltau)
        self.precon_csr = csr.to(self.device)
        self.precon_table = tab.to(self.device)

    def _ptab(self):
        if not (self.dtau <= 0.1 and self.precon_on):
            return None
        if self.precon_table is None:
            self.reset_precon()
        return self.precon_table

    # --------------------------------------------------------------- layouts
    def _perm(self, boson):
        """[bs, 2, Lx, Ly, Ltau] -> contiguous [bs, Ltau*Ly*Lx*2]."""
        return boson.permute([0, 4, 3, 2, 1]).reshape(boson.shape[0], -1).contiguous()

    def _perm_copy(self, t):
        """Fresh contiguous float64 copy in the kernels' layout (never aliases t, which may be a
        permuted view whose permutation back is already contiguous)."""
        out = torch.empty(t.shape[0], t[0].numel(), dtype=torch.float64, device=t.device)
        out.copy_(t.permute([0, 4, 3, 2, 1]).reshape(t.shape[0], -1))
        return out

    def _unperm(self, flat, bs=None):
        bs = flat.shape[0] if bs is None else bs
        return flat.view(bs, self.Ltau, self.Ly, self.Lx, 2).permute(0, 4, 3, 2, 1)

    def _phi_for(self, boson_flat, vec_dtype):
        want = torch.float32 if vec_dtype == torch.complex64 else torch.float64
        return boson_flat.to(want).contiguous()

    # ------------------------------------------------------------- draws
    def draw_momentum(self):
        """[bs, 2, Lx, Ly, Ltau] ~ N(0, m) (hmc_sampler_batch.py:1247-1252)."""
        return torch.randn(self.bs, 2, self.Lx, self.Ly, self.Ltau, device=self.device, dtype=self.dtype) * math.sqrt(self.m)

    # ------------------------------------------- Fourier-accelerated momenta
    # The reference's adaptive mass (fourier_accel.py): estimated when mass_mode != 0, never
    # used by its trajectory, which draws draw_momentum() (hmc_sampler_batch.py:2389-2390).
    sigma_hat = property(lambda self: self._fa.sigma_hat, lambda self, v: setattr(self._fa, "sigma_hat", v))
    sigma_hat_cpu = property(lambda self: self._fa.sigma_hat_cpu, lambda self, v: setattr(self._fa, "sigma_hat_cpu", v))
    boson_mean_cpu = property(lambda self: self._fa.boson_mean_cpu)
    multiplier = property(lambda self: self._fa.multiplier)

    def update_sigma_hat_cpu(self, boson_cpu, i):
        """hmc_sampler_batch.py:1202-1210."""
        fourier_accel.update_sigma_hat(self._fa, boson_cpu, i)

    def apply_sigma_hat_cpu(self, i):
        """hmc_sampler_batch.py:1133-1200."""
        fourier_accel.apply_sigma_hat(self._fa, i, self.mass_mode)

    def stabilize_sigma_hat(self, method="shrink", **kwargs):
        """hmc_sampler_batch.py:1212-1245."""
        self._fa.sigma_hat = fourier_accel.stabilize_sigma_hat(self._fa.sigma_hat, method, **kwargs)

    def draw_momentum_fft(self):
        """[bs, 2, Lx, Ly, Ltau] ~ N(0, F^-1 diag(1/sigma_hat) F) (hmc_sampler_batch.py:1254-1265)."""
        return fourier_accel.draw_momentum_fft(self._fa)

    def apply_m_inv(self, p):
        """M^-1 p = F^-1 sigma_hat F p (hmc_sampler_batch.py:1267-1277)."""
        return fourier_accel.apply_m_inv(self._fa, p)

    def draw_psudo_fermion(self):
        """randn(complex)/sqrt(2), [bs, Ltau*Ly*Lx] (hmc_sampler_batch.py:1279-1287)."""
        return torch.randn(self.bs, self.Lx * self.Ly * self.Ltau, device=self.device, dtype=self.cdtype) / math.sqrt(2)

    # ----------------------------------------------------------- boson sector
    def _boson_parts(self, x):
        cJ = 1.0 / (self.J * self.dtau ** 2)
        _, parts = kernels.hamiltonian(self._perm(x).double(), self.Lx, self.Ltau, cJ, self.K, self.compact)
        return parts

    def _boson_forces(self, x, compact):
        cJ = 1.0 / (self.J * self.dtau ** 2)
        Fp, Ft = kernels.boson_force(self._perm(x).double(), self.Lx, self.Ltau, cJ, self.K, compact)
        return self._unperm(Fp).to(x.dtype), self._unperm(Ft).to(x.dtype)

    def action_boson_tau(self, x):
        """cJ * sum (1 - cos) or cJ * sum d^2/2 (hmc_sampler_batch.py:1363-1372), per chain."""
        return self._boson_parts(x)[:, 0].to(x.dtype)

    def action_boson_plaq_matfree(self, boson):
        """K * sum cos(curl) (hmc_sampler_batch.py:1428-1436), per chain."""
        return self._boson_parts(boson)[:, 1].to(boson.dtype)

    def force_b_plaq_matfree(self, boson):
        return self._boson_forces(boson, self.compact)[0]

    def force_b_tau_cmp(self, boson):
        return self._boson_forces(boson, True)[1]

    def force_b_tau_noncmp(self, boson):
        return self._boson_forces(boson, False)[1]

    def curl_phi(self, phi):
        """phi_x(r) + phi_y(r+x) - phi_x(r+y) - phi_y(r) (hmc_sampler_batch.py:1412-1426)."""
        return phi[:, 0] + torch.roll(phi[:, 1], -1, 1) - torch.roll(phi[:, 0], -1, 2) - phi[:, 1]

    # ----------------------------------------------------------------- solver
    def preconditioned_cg_fast_test(self, boson, b, MhM_inv=None, matL=None, b_idx=None, axs=None, cg_dtype=None, MhM=None):
        """Batched PCG for M'M x = b (hmc_sampler_batch.py:751-846).

        boson: [bs, Ltau, Ly*Lx*2] (permuted) or [bs, Ltau*Ly*Lx*2]; b: [bs, Ltau*Ly*Lx].
        Returns (x, iterations, error): iterations is None under CUDA graphs like the reference."""
        bs = b.shape[0]
        phi = boson.reshape(bs, -1)
        ptab = self._ptab()
        status = torch.empty(bs, dtype=torch.int32, device=b.device)
        if self.precision == "mixed":
            x, iters, err = kernels.pcg_mixed_solve(phi.to(torch.float64).contiguous(), b.to(torch.complex128).contiguous(),
                                                    self.Lx, self.Ltau, self.dtau, ptab=ptab, rtol=self.cg_rtol,
                                                    max_iter=self.max_iter, ru_period=self.ru_period, status=status)
        else:
            x, iters, err = kernels.pcg_solve(self._phi_for(phi, b.dtype), b.contiguous(), self.Lx, self.Ltau, self.dtau,
                                              ptab=ptab, rtol=self.cg_rtol, max_iter=self.max_iter, status=status)
        self.last_solve_status = status
        if not self.cuda_graph and not torch.cuda.is_current_stream_capturing():   # readback: outside capture only
            codes = status.tolist()
            if max(codes) >= 2:
                warnings.warn(f"preconditioned_cg_fast_test: chain statuses {codes} (2: non-finite, 3: breakdown); "
                              "those chains stopped at their last finite iterate", RuntimeWarning)
        iterations = None if self.cuda_graph else iters.to(torch.int64)
        return x, iterations, err.to(self.dtype)

    def Ot_inv_psi_fast(self, psi, boson, MhM):
        """(M'M)^{-1} psi (hmc_sampler_batch.py:2112-2131)."""
        return self.preconditioned_cg_fast_test(boson, psi, MhM_inv=self.precon, MhM=MhM)

    def force_f_fast(self, psi, boson, MhM):
        """(Ft, xi, iters, r_err) with Ft = -dS_f/dphi, [bs, 2, Lx, Ly, Ltau] (hmc_sampler_batch.py:2150-2227).

        The per-(b, tau) loop of b_vec_per_tau calls and gathers is one fused adjoint kernel."""
        bs = psi.shape[0]
        phi = self._perm(boson)
        xi, iters, err = self.Ot_inv_psi_fast(psi, phi, MhM)
        F = kernels.fermion_force(self._phi_for(phi, xi.dtype), xi, self.Lx, self.Ltau, self.dtau)
        Ft = self._unperm(F, bs).to(self.dtype)
        return Ft, xi.view(bs, -1), iters, err

    # ---------------------------------------------------------------- leapfrog
    def leapfrog_cmp(self, x, p, dt, tau_mask, force_b_plaq, force_b_tau):
        """5 boson sub-steps of dt/5 (hmc_sampler_batch.py:2332-2347); tensors in the reference layout."""
        cJ = 1.0 / (self.J * self.dtau ** 2)
        xf, pf = self._perm_copy(x), self._perm_copy(p)
        Fp, Ft = self._perm_copy(force_b_plaq), self._perm_copy(force_b_tau)
        dtv = dt.reshape(-1).to(torch.float64).expand(x.shape[0]).contiguous()
        mask = tau_mask.reshape(-1).to(torch.float64).contiguous() if tau_mask is not None else None
        kernels.leapfrog_boson(xf, pf, Fp, Ft, dtv, self.Lx, self.Ltau, cJ, self.K, self.compact, m=self.m, nsub=5,
                               tau_mask=mask)
        u = lambda t: self._unperm(t).to(x.dtype)
        return u(xf), u(pf), u(Fp), u(Ft)

    def _get_engine(self):
        if self._engine is None:
            J_in = self.J * self.dtau / self.Nf * 4
            K_in = self.K / self.dtau / self.Nf * 2
            self._engine = TrajectoryEngine(self.Lx, self.Ltau, self.bs, dtau=self.dtau, J=J_in, K=K_in, compact=self.compact,
                                            m=self.m, delta_t=self.delta_t, n_leap=self.N_leapfrog, nsub=5,
                                            precision=self.precision, max_iter=self.max_iter, rtol=self.cg_rtol,
                                            ru_period=self.ru_period, precon=self._ptab() is not None, device=self.device)
        return self._engine

    def leapfrog_proposer5(self, boson, tau_mask):
        """One multi-timescale leapfrog trajectory (hmc_sampler_batch.py:2366-2625).

        Returns (x, H0, H_fin, mean cg iterations, mean cg r_err) with x in the reference layout."""
        eng = self._get_engine()
        eng.set_phi(self._perm(boson).double())
        eng.dt.copy_(self.delta_t_tensor.to(torch.float64))
        eng.tau_mask.copy_(tau_mask.reshape(-1).to(torch.float64))
        if self.cuda_graph:
            if eng.graph is None:
                eng.capture(n_warmup=1, accept=False, pool=self.graph_memory_pool)
                self.graph_memory_pool = eng.graph.pool()
                eng.set_phi(self._perm(boson).double())
            eng.replay()
        else:
            eng.propose()
        x = self._unperm(eng.x.clone()).to(boson.dtype)
        self.last_solve_status = eng.status.clone()
        return x, eng.H0.clone(), eng.H1.clone(), eng.iters.mean(dim=0), eng.err.mean(dim=0)

    def metropolis_update(self):
        """One Metropolis step per tau block: propose with the block's tau mask, accept with
        rand < exp(H_old - H_new), record thresholds (hmc_sampler_batch.py:2886-2930).  With
        several CG gears, the first applies after gear0_steps sweeps (2896-2899)."""
        gears = self._MAX_ITERS_TO_CAPTURE
        gear0_steps = _env("gear0_steps", 1000, int)
        want = gears[0] if len(gears) == 1 or self.cur_step > gear0_steps else gears[1]
        if want != self.max_iter:
            self.max_iter = want
            self._engine = None
        accp = cg_converge_iter = cg_r_err = None
        while self.tau_block_idx < self.max_tau_block_idx:
            t0 = self.tau_block_idx * self.tau_block_size
            tau_mask = torch.zeros((1, 1, 1, 1, self.Ltau), device=self.device, dtype=self.dtype)
            tau_mask[..., t0:t0 + self.tau_block_size] = 1.0
            boson_new, H_old, H_new, cg_converge_iter, cg_r_err = self.leapfrog_proposer5(self.boson, tau_mask)
            accp = torch.rand(self.bs, device=self.device, dtype=torch.float64) < torch.exp(H_old - H_new)
            # a chain whose solve failed (non-finite or breakdown) is rejected and reported
            failed = (self.last_solve_status >= 2).any(dim=0)
            accp &= ~failed
            self.boson[accp] = boson_new[accp]
            thresholds = torch.minimum(torch.exp(H_old - H_new), torch.ones_like(H_old))
            thresholds = torch.where(failed, torch.zeros_like(thresholds), thresholds)
            for b, th in enumerate(thresholds.tolist()):   # one device-to-host copy per sweep
                self.threshold_queue[b].append(th)
            nfail = int(failed.sum().item())
            if nfail:
                warnings.warn(f"metropolis_update: {nfail} of {self.bs} chains rejected: a PCG solve was non-finite "
                              f"or broke down (status {self.last_solve_status.max().item()})", RuntimeWarning)
            self.tau_block_idx += 1
        self.tau_block_idx = 0
        return self.boson, accp, cg_converge_iter, cg_r_err

    def adjust_delta_t(self):
        """Per-chain step size x0.9 / x1.1 outside the acceptance window (hmc_sampler_batch.py:699-724)."""
        adjust_delta_t(self.delta_t_tensor, self.threshold_queue, self.lower_limit, self.upper_limit)

    # ------------------------------------------------------------------ graphs
    def initialize_force_graph(self):
        """ForceGraphRunner per max_iter (hmc_sampler_batch.py:336-366)."""
        from .force_graph_runner import ForceGraphRunner
        if not self.cuda_graph:
            return
        dummy_psi = torch.zeros(self.bs, self.Vs * self.Ltau, dtype=self.cdtype, device=self.device)
        dummy_boson = torch.zeros(self.bs, 2, self.Lx, self.Ly, self.Ltau, dtype=self.dtype, device=self.device)
        runner = ForceGraphRunner(self)
        self.graph_memory_pool = runner.capture(dummy_psi, dummy_boson, self.max_iter, self.graph_memory_pool)
        self.force_graph_runners[self.max_iter] = runner

    def initialize_leapfrog_cmp_graph(self):
        """LeapfrogCmpGraphRunner (hmc_sampler_batch.py:403-440)."""
        from .metropolis_graph_runner import LeapfrogCmpGraphRunner
        if not self.cuda_graph:
            return
        z = lambda: torch.zeros(self.bs, 2, self.Lx, self.Ly, self.Ltau, dtype=self.dtype, device=self.device)
        runner = LeapfrogCmpGraphRunner(self)
        self.graph_memory_pool = runner.capture(z(), z(), torch.zeros(self.bs, 1, 1, 1, 1, dtype=self.dtype, device=self.device),
                                                torch.ones((1, 1, 1, 1, self.Ltau), dtype=self.dtype, device=self.device),
                                                z(), z(), self.graph_memory_pool)
        self.leapfrog_cmp_graph_runners = runner

    def couplings(self):
        """(cJ, cK) used by the kernels."""
        return 1.0 / (self.J * self.dtau ** 2), self.K


__all__ = ["HmcSampler", "couplings"]
