"""Device-resident HMC trajectory for a batch of chains (leapfrog_proposer5 +
metropolis_update, qed_fermion/hmc_sampler_batch.py:2366-2625, 2886-2930).

Everything stays in HBM between steps: the boson field phi (float64 master in
the fermion-side layout [chain][tau][y][x][mu], plus a float32 copy for the
FP32 operator), the momenta, the pseudofermion and the solution.  One
trajectory is

    p ~ N(0, m);  R ~ randn(complex)/sqrt(2);  psi = M^dag(phi) R
    xi = (M^dag M)^{-1} psi;  Ff = -dS_f/dphi
    H0 = S_tau + S_plaq + p^2/2m + Re(psi^dag xi)
    N_leapfrog times: p += dt/2 Ff; 5 boson sub-steps; solve; force; p += dt/2 Ff
    H1 = ...;  accept where rand < exp(H0 - H1)

with every piece a HIP kernel of libqf_hip.so or a torch elementwise op on the
same stream, so the whole trajectory is captured into one hipGraph
(torch.cuda.graph) and replayed with a single launch from the host.

precision:
  "fp64"  complex128 PCG, float64 operator and force (the plain path)
  "mixed" complex64 PCG with FP64 dot products and FP64 reliable updates every
          ru_period iterations (true residual from the float64 operator);
          psi, xi, force and H in FP64
  "fp32"  complex64 PCG and force on the float32 field (the reference's dtypes)
"""
import math

import torch

from . import _lib, kernels


def adjust_delta_t(delta_t, queues, lower_limit=0.25, upper_limit=0.5):
    """Per-chain step-size control of hmc_sampler_batch.py:699-724, in place on ``delta_t``
    (a [bs] tensor, any device): a chain whose deque of clipped thresholds min(1, e^{-dH}) is
    full and averages below lower_limit gets dt x 0.9, above upper_limit dt x 1.1; its deque is
    then cleared.  Chains with a partial deque, or an average inside the window, are left alone.
    Returns the list of chains adjusted."""
    changed = []
    for b, q in enumerate(queues):
        if len(q) < q.maxlen:
            continue
        avg = sum(q) / len(q)
        if avg < lower_limit:
            delta_t[b] *= 0.9
        elif upper_limit < avg:
            delta_t[b] *= 1.1
        else:
            continue
        q.clear()
        changed.append(b)
    return changed


def couplings(J, K, dtau, Nf=2):
    """(cJ, cK) = (1/(J' dtau^2), K') with J' = J/dtau*Nf/4, K' = K*dtau*Nf/2 (hmc_sampler_batch.py:173-174, 1344)."""
    Jp = J / dtau * Nf / 4
    return 1.0 / (Jp * dtau ** 2), K * dtau * Nf / 2


class TrajectoryEngine:
    def __init__(self, L, Ltau, bs, *, dtau=0.1, J=1.0, K=1.0, compact=False, m=1.0, delta_t=0.06, n_leap=5, nsub=5,
                 precision="mixed", max_iter=400, rtol=1e-9, ru_period=50, precon=True, device="cuda"):
        if precision not in ("fp64", "mixed", "fp32"):
            raise ValueError(f"precision {precision!r}")
        self.L, self.Ltau, self.bs = L, Ltau, bs
        self.N = L * L * Ltau
        self.dtau, self.compact, self.m = dtau, bool(compact), m
        self.cJ, self.cK = couplings(J, K, dtau)
        self.n_leap, self.nsub = n_leap, nsub
        self.precision, self.rtol, self.ru_period = precision, rtol, ru_period
        self.max_iter = max_iter
        dev = torch.device(device)
        self.device = dev
        f64, c128 = torch.float64, torch.complex128
        cvec = torch.complex64 if precision == "fp32" else c128
        N2 = 2 * self.N
        self.x = torch.zeros(bs, N2, dtype=f64, device=dev)
        self.x0 = torch.zeros_like(self.x)
        self.x32 = torch.zeros(bs, N2, dtype=torch.float32, device=dev)
        self.p = torch.zeros_like(self.x)
        self.Fp = torch.zeros_like(self.x)
        self.Ft = torch.zeros_like(self.x)
        self.Ff = torch.zeros_like(self.x)
        self.Ff32 = torch.zeros(bs, N2, dtype=torch.float32, device=dev) if precision == "fp32" else None
        self.R = torch.zeros(bs, self.N, dtype=cvec, device=dev)
        self.psi = torch.zeros_like(self.R)
        self.xi = torch.zeros_like(self.R)
        self.dt = torch.full((bs,), float(delta_t), dtype=f64, device=dev)
        self.tau_mask = torch.ones(Ltau, dtype=f64, device=dev)
        self.H0 = torch.zeros(bs, dtype=f64, device=dev)
        self.H1 = torch.zeros_like(self.H0)
        self.parts0 = torch.zeros(bs, 4, dtype=f64, device=dev)
        self.parts1 = torch.zeros_like(self.parts0)
        self.err = torch.zeros(n_leap + 1, bs, dtype=f64, device=dev)
        self.iters = torch.zeros_like(self.err)
        # QF_SOLVE_* code of every solve of the last trajectory, per chain
        self.status = torch.zeros(n_leap + 1, bs, dtype=torch.int32, device=dev)
        self.accept = torch.zeros(bs, dtype=torch.bool, device=dev)
        self.u = torch.zeros(bs, dtype=f64, device=dev)
        if precon:
            self.ptab = kernels.precon_table(Ltau, dtype=torch.float32 if precision != "fp64" else f64, device=dev)
        else:
            self.ptab = None
        prec = _lib.QF_PREC_F64 if precision == "fp64" else _lib.QF_PREC_F32
        nbytes = kernels.pcg_workspace_bytes(prec, precision == "mixed", bs, L, Ltau)
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.graph = None

    # ----------------------------------------------------------- iteration caps
    @property
    def max_iter(self):
        """The largest PCG iteration cap of the trajectory's solves."""
        return max(self.max_iters)

    @max_iter.setter
    def max_iter(self, v):
        """One cap for every solve (an int), or one per solve: solve 0 is the pseudofermion's,
        solve k the one after leapfrog step k.  Caps are baked into a captured graph: set
        them before capture().  The reference captures one graph per cap
        (_MAX_ITERS_TO_CAPTURE, hmc_sampler_batch.py:305-309)."""
        caps = [int(v)] * (self.n_leap + 1) if isinstance(v, (int, float)) else [int(c) for c in v]
        if len(caps) != self.n_leap + 1 or min(caps) < 1:
            raise ValueError(f"need {self.n_leap + 1} positive iteration caps, got {caps}")
        self.max_iters = caps

    # ------------------------------------------------------------------ state
    def set_phi(self, phi):
        """phi: [bs, Ltau*L*L*2] (fermion layout) in any real dtype."""
        self.x.copy_(phi.reshape(self.bs, -1))

    # --------------------------------------------------------------- pieces
    def _solve(self, k):
        L, Lt = self.L, self.Ltau
        if self.precision == "mixed":
            kernels.pcg_mixed_solve(self.x, self.psi, L, Lt, self.dtau, ptab=self.ptab, rtol=self.rtol,
                                    max_iter=self.max_iters[k], ru_period=self.ru_period, phi32=self.x32, x=self.xi,
                                    work=self.work, err=self.err[k], iters=self.iters[k], status=self.status[k])
        elif self.precision == "fp64":
            kernels.pcg_solve(self.x, self.psi, L, Lt, self.dtau, ptab=self.ptab, rtol=self.rtol, max_iter=self.max_iters[k],
                              x=self.xi, work=self.work, err=self.err[k], iters=self.iters[k], status=self.status[k])
        else:
            kernels.pcg_solve(self.x32, self.psi, L, Lt, self.dtau, ptab=self.ptab, rtol=self.rtol,
                              max_iter=self.max_iters[k], x=self.xi, work=self.work, err=self.err[k], iters=self.iters[k],
                              status=self.status[k])

    def _force(self):
        if self.precision == "fp32":
            kernels.fermion_force(self.x32, self.xi, self.L, self.Ltau, self.dtau, out=self.Ff32)
            self.Ff.copy_(self.Ff32)
        else:
            kernels.fermion_force(self.x, self.xi, self.L, self.Ltau, self.dtau, out=self.Ff)

    def _ham(self, H, parts):
        kernels.hamiltonian(self.x, self.L, self.Ltau, self.cJ, self.cK, self.compact, self.m, p=self.p, psi=self.psi,
                            xi=self.xi, H=H, parts=parts)

    def propose(self, draw=True):
        """One leapfrog_proposer5 trajectory from self.x; leaves the proposal in self.x, the start in self.x0."""
        L, Lt = self.L, self.Ltau
        self.x0.copy_(self.x)
        if draw:
            self.p.normal_()
            self.p.mul_(math.sqrt(self.m))
            self.R.copy_(torch.randn(self.R.shape, dtype=self.R.dtype, device=self.device) / math.sqrt(2))
        if self.precision != "fp64":
            self.x32.copy_(self.x)
        phi_op = self.x32 if self.precision == "fp32" else self.x
        kernels.apply("Mdag", phi_op, self.R, L, Lt, self.dtau, out=self.psi)
        self._solve(0)
        self._force()
        self._ham(self.H0, self.parts0)
        kernels.boson_force(self.x, L, Lt, self.cJ, self.cK, self.compact, Fp=self.Fp, Ft=self.Ft)
        for leap in range(self.n_leap):
            kernels.momentum_kick(self.p, self.Ff, self.dt, 0.5, L, Lt, tau_mask=self.tau_mask)
            kernels.leapfrog_boson(self.x, self.p, self.Fp, self.Ft, self.dt, L, Lt, self.cJ, self.cK, self.compact,
                                   m=self.m, nsub=self.nsub, tau_mask=self.tau_mask)
            if self.precision != "fp64":
                self.x32.copy_(self.x)
            self._solve(leap + 1)
            self._force()
            kernels.momentum_kick(self.p, self.Ff, self.dt, 0.5, L, Lt, tau_mask=self.tau_mask)
        self._ham(self.H1, self.parts1)

    def solve_failed(self):
        """Per chain: True if a solve of the last trajectory ended non-finite or in a breakdown."""
        return (self.status >= _lib.QF_SOLVE_NONFINITE).any(dim=0)

    def metropolis(self):
        """accp = rand < exp(H_old - H_new); rejected chains return to x0 (hmc_sampler_batch.py:2910-2918).
        A chain whose solve failed (non-finite or breakdown) is rejected, on the device."""
        self.u.uniform_()
        torch.lt(self.u, torch.exp(self.H0 - self.H1), out=self.accept)
        self.accept.logical_and_(~self.solve_failed())
        self.x.copy_(torch.where(self.accept.view(-1, 1), self.x, self.x0))

    def trajectory(self):
        self.propose()
        self.metropolis()

    # ---------------------------------------------------------------- graph
    # A captured graph holds every PCG iteration up to its cap (two kernels each, plus the
    # reliable-update checks); ending the capture of a graph of about 330,000 kernel nodes
    # crashed the HIP runtime on the host, 120,000 captured fine.
    MAX_GRAPH_ITERATIONS = 65536

    def capture(self, n_warmup=1, pool=None, draw=True, accept=True):
        """Capture one trajectory (draws, solves, forces, leapfrog, accept) as a hipGraph.
        draw=False leaves p and R to the caller; accept=False stops at the proposal.  The
        solves' caps together may hold at most MAX_GRAPH_ITERATIONS iterations: choose caps per
        solve (max_iter as a list) from the iterations the solves need."""
        if sum(self.max_iters) > self.MAX_GRAPH_ITERATIONS:
            raise ValueError(f"a trajectory graph of {sum(self.max_iters)} PCG iterations (caps {self.max_iters}) exceeds "
                             f"{self.MAX_GRAPH_ITERATIONS}; lower the caps per solve")
        def body():
            self.propose(draw=draw)
            if accept:
                self.metropolis()

        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        saved = [t.clone() for t in (self.x, self.p, self.R)]
        with torch.cuda.stream(s):
            for _ in range(n_warmup):
                body()
        torch.cuda.current_stream().wait_stream(s)
        for t, v in zip((self.x, self.p, self.R), saved):
            t.copy_(v)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, pool=pool):
            body()
        self.graph = g
        return g.pool()

    def replay(self):
        self.graph.replay()

    def acceptance(self):
        """Clipped Metropolis threshold min(1, e^{H0 - H1}) per chain (hmc_sampler_batch.py:2921)."""
        return torch.minimum(torch.exp(self.H0 - self.H1), torch.ones_like(self.H0))
