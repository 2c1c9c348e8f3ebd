"""Bond-bond stochastic estimators (qed_fermion/stochastic_estimator.py:1393-1654, 2594-2651).

``BondBondEstimator`` computes T1, T21, T2, T4 (pairs of noise vectors s < s')
and Tvv / Tv (single noise vectors) from eta, G_eta = M^{-1} eta and the
chain's boson field, and combines them like ``get_bond_bond_per_b2``.

Every estimator is C(Delta) = (1/V) sum_r a(r - Delta) b(r), averaged over
pairs and tau.  The reference evaluates it per batch of int(0.1 Nrv) pairs
with three FFTs and a host loop (1405-1432).  Here the k-space sum
S(k) = sum over pairs and slices of Ahat(k) Bhat(-k) is formed in FP64 and
C(Delta) = fft2(S)(Delta) / V^2 / (pairs * Ltau): one transform of an [L, L]
array per estimator.  Two ways to form S:

* "fused" (default where the slices fit in LDS): one launch of
  ``qf_obs_fused`` per estimator; each work-group builds a and b of one
  (slice, pair) in LDS, transforms them with its own mixed-radix FFT and
  accumulates in registers.  No intermediate touches HBM.
* "rocfft": per chunk of pairs (sized to a memory budget) ``qf_obs_fields``
  (HIP) writes a and b, torch.fft (rocFFT) transforms them and
  ``qf_obs_kspace`` (HIP) adds the products into FP64 partials.

The partials are reduced in a fixed order (``qf_obs_reduce``).  The same-noise-vector terms each estimator adds
at one Delta come from ``qf_obs_diag``.  Inputs stay where they are (no
rolled copies); all buffers are allocated once per shape, so a call can be
captured in a graph after one warm-up call.
"""
import torch

from . import _lib
from ._lib import QfError, call, ptr, stream_handle

_NAMES = {"T1": _lib.QF_OBS_T1, "T21": _lib.QF_OBS_T21, "T2": _lib.QF_OBS_T2, "T4": _lib.QF_OBS_T4}
# Delta at which each estimator adds its diagonal term, and which of qf_obs_diag's three sums
_DIAG = {"T1": (lambda L: (0, L - 1), 0), "T21": (lambda L: (0, 0), 1), "T2": (lambda L: (0, 0), 1), "T4": (lambda L: (0, 1), 2)}
_PREC = {torch.complex64: _lib.QF_PREC_F32, torch.complex128: _lib.QF_PREC_F64}
_REAL = {torch.complex64: torch.float32, torch.complex128: torch.float64}


def reorder_fft_grid2(t, dims=(-2, -1)):
    """FFT order -> ascending momentum order on two axes (stochastic_estimator.py:253-256)."""
    return torch.roll(t, shifts=(t.shape[dims[0]] // 2, t.shape[dims[1]] // 2), dims=dims)


def reorder_fft_grid2_inverse(t, dims=(-2, -1)):
    """Inverse of reorder_fft_grid2 (stochastic_estimator.py:259-262)."""
    return torch.roll(t, shifts=(-t.shape[dims[0]] // 2, -t.shape[dims[1]] // 2), dims=dims)


class BondBondEstimator:
    """Workspaces and launches for one (L, Ltau, Nrv, dtype, device).

    pair_chunk: pairs per chunk (default: as many as fit ``chunk_bytes`` per
    field buffer; two such buffers are held, and the transforms of a chunk take two more).  nwg: work-groups of the
    k-space accumulation (its FP64 partials are [nwg, L*L])."""

    def __init__(self, L, Ltau, nrv, dtype=torch.complex64, device="cuda", pair_chunk=None, chunk_bytes=256 << 20, nwg=512,
                 method="auto"):
        if dtype not in _PREC:
            raise QfError(_lib.QF_ERR_INVALID, f"unsupported dtype {dtype}")
        if nrv < 2:
            raise QfError(_lib.QF_ERR_INVALID, "the pair estimators need at least two noise vectors")
        if L < 2:
            raise QfError(_lib.QF_ERR_INVALID, f"L = {L}")
        self.L, self.Ltau, self.nrv, self.dtype = L, Ltau, nrv, dtype
        self.device = torch.device(device)
        self.V = L * L
        self.N = self.V * Ltau
        self.npairs = nrv * (nrv - 1) // 2
        esize = torch.empty((), dtype=dtype).element_size()
        if pair_chunk is None:
            pair_chunk = max(1, chunk_bytes // (self.N * esize))
        self.chunk = int(min(pair_chunk, max(self.npairs, nrv)))
        if method not in ("auto", "fused", "rocfft"):
            raise ValueError(f"method {method!r}")
        dev = self.device
        fused_nwg = 0
        if method != "rocfft":
            with torch.cuda.device(dev):
                fused_nwg = _lib.lib().qf_obs_fused_grid(_PREC[dtype], L)
            if method == "fused" and fused_nwg <= 0:
                raise QfError(_lib.QF_ERR_UNSUPPORTED, f"no fused estimator kernel for L = {L}, {dtype}")
        self.method = "fused" if fused_nwg > 0 else "rocfft"
        self.nwg = int(fused_nwg if self.method == "fused" else nwg)
        self.nblk = 256
        self.pairs = torch.combinations(torch.arange(nrv), r=2).to(torch.int32).reshape(-1).contiguous().to(dev)
        if self.method == "rocfft":
            self.A = torch.empty(self.chunk, self.N, dtype=dtype, device=dev)
            self.B = torch.empty_like(self.A)
        self.part = torch.empty(2 * self.nwg, self.V, dtype=torch.complex128, device=dev)
        self.S = torch.empty(2, self.V, dtype=torch.complex128, device=dev)
        self.dpart = torch.empty(self.nblk, 3, dtype=torch.complex128, device=dev)
        self.dsum = torch.empty(3, dtype=torch.complex128, device=dev)
        self.U = torch.empty(self.N, dtype=dtype, device=dev)          # e^{i phix} of the current call

    # ------------------------------------------------------------ helpers
    def _inputs(self, eta, G_eta, boson):
        L, Lt = self.L, self.Ltau
        if eta.dtype != self.dtype or G_eta.dtype != self.dtype:
            raise QfError(_lib.QF_ERR_INVALID, f"eta / G_eta must be {self.dtype}")
        if eta.numel() != self.nrv * self.N or G_eta.numel() != self.nrv * self.N:
            raise QfError(_lib.QF_ERR_INVALID, "eta / G_eta sizes do not match [Nrv, Ltau*L*L]")
        if tuple(boson.shape[-4:]) != (2, L, L, Lt):
            raise QfError(_lib.QF_ERR_INVALID, f"boson must be [1, 2, {L}, {L}, {Lt}]")
        for t in (eta, G_eta, boson):
            if not t.is_cuda:
                raise QfError(_lib.QF_ERR_INVALID, "tensor is not on the GPU; the HIP path has no CPU fallback")
        phix = boson.reshape(-1, 2, L, L, Lt)[0, 0].permute(2, 1, 0).to(_REAL[self.dtype]).contiguous()   # [tau, y, x]
        call("qf_obs_phase", _PREC[self.dtype], ptr(phix), ptr(self.U), self.N, stream_handle())
        return eta.contiguous(), G_eta.contiguous(), phix

    def _fft(self, src, n):
        """Forward 2-D transforms of the first n fields (torch's rocFFT plan; writing into a given
        out= buffer would add a full copy of the result)."""
        L = self.L
        return torch.fft.fft2(src[:n].view(n * self.Ltau, L, L))

    def _kspace_sum(self, obs, eta, G, phix, count):
        """S[0] = sum_k-products over all pairs (or noise vectors for OBS_V), S[1] = sum of Ahat (OBS_V)."""
        prec = _PREC[self.dtype]
        lin = 1 if obs == _lib.QF_OBS_V else 0
        st = stream_handle()
        if self.method == "fused":
            call("qf_obs_fused", prec, obs, ptr(eta), ptr(G), ptr(self.U), ptr(self.pairs), count, self.nrv, self.L, self.Ltau,
                 ptr(self.part), self.nwg, st)
            call("qf_obs_reduce", ptr(self.part), self.nwg, self.V, ptr(self.S[0]), st)
            if lin:
                call("qf_obs_reduce", ptr(self.part[self.nwg:]), self.nwg, self.V, ptr(self.S[1]), st)
            return
        first = True
        for p0 in range(0, count, self.chunk):
            n = min(self.chunk, count - p0)
            call("qf_obs_fields", prec, obs, ptr(eta), ptr(G), ptr(self.U), ptr(self.pairs), p0, n, self.nrv, self.L, self.Ltau,
                 ptr(self.A), ptr(self.B), st)
            ah = self._fft(self.A, n)
            bh = ah if lin else self._fft(self.B, n)
            call("qf_obs_kspace", prec, ptr(ah), ptr(bh), n * self.Ltau, self.L, lin, 0 if first else 1, ptr(self.part),
                 self.nwg, st)
            first = False
        call("qf_obs_reduce", ptr(self.part), self.nwg, self.V, ptr(self.S[0]), st)
        if lin:
            call("qf_obs_reduce", ptr(self.part[self.nwg:]), self.nwg, self.V, ptr(self.S[1]), st)

    def _corr(self, S):
        """fft2(S) / V^2 over [L, L]: (1/V) sum_r a(r - Delta) b(r) summed over the slices of S."""
        return torch.fft.fft2(S.view(self.L, self.L)) / (self.V * self.V)

    def _diag(self, eta, G, phix):
        call("qf_obs_diag", _PREC[self.dtype], ptr(eta), ptr(G), ptr(self.U), self.nrv, self.L, self.Ltau, ptr(self.dpart), self.nblk,
             stream_handle())
        call("qf_obs_reduce", ptr(self.dpart), self.nblk, 3, ptr(self.dsum), stream_handle())
        return self.dsum / (self.nrv * self.N)

    # ------------------------------------------------------------ estimators
    def T(self, name, eta, G_eta, boson, diag=None):
        """T1 / T21 / T2 / T4 (stochastic_estimator.py:1393-1610): [Ly, Lx] in eta's dtype.
        diag: the three same-noise-vector sums of these inputs when already computed."""
        obs = _NAMES[name]
        eta, G, phix = self._inputs(eta, G_eta, boson)
        self._kspace_sum(obs, eta, G, phix, self.npairs)
        out = -self._corr(self.S[0]) / (self.npairs * self.Ltau)
        (iy, ix), which = _DIAG[name][0](self.L), _DIAG[name][1]
        d = self._diag(eta, G, phix) if diag is None else diag
        out[iy, ix] += d[which]
        return out.to(self.dtype)

    def diag_sums(self, eta, G_eta, boson):
        """The three same-noise-vector sums the pair estimators add at one Delta each."""
        eta, G, phix = self._inputs(eta, G_eta, boson)
        return self._diag(eta, G, phix).clone()

    def Tvv(self, eta, G_eta, boson):
        """(Tvv, Tv) (stochastic_estimator.py:1612-1654): [Ly, Lx] each, in eta's dtype."""
        eta, G, phix = self._inputs(eta, G_eta, boson)
        self._kspace_sum(_lib.QF_OBS_V, eta, G, phix, self.nrv)
        tvv = self._corr(self.S[0]) / (self.nrv * self.Ltau)
        tv = torch.fft.ifft2(self.S[1].view(self.L, self.L)) / (self.nrv * self.Ltau)
        return tvv.to(self.dtype), tv.to(self.dtype)

    def bond_bond(self, eta, G_eta, boson, runners=None):
        """get_bond_bond_per_b2 (2594-2641): {'BB0_r', 'BB_r', 'B_r', 'BB_k'}, real [Ly, Lx].
        runners: optional {name: callable(eta, G_eta, boson)} replacing the direct calls (graph runners)."""
        f = runners or {}
        d = None if runners else self.diag_sums(eta, G_eta, boson)    # shared by the four pair estimators
        t1 = f.get("T1", lambda *a: self.T("T1", *a, diag=d))(eta, G_eta, boson)
        t21 = f.get("T21", lambda *a: self.T("T21", *a, diag=d))(eta, G_eta, boson)
        t2 = f.get("T2", lambda *a: self.T("T2", *a, diag=d))(eta, G_eta, boson)
        t4 = f.get("T4", lambda *a: self.T("T4", *a, diag=d))(eta, G_eta, boson)
        tvv, tv = f.get("Tvv", self.Tvv)(eta, G_eta, boson)
        BB_r = (t1 + t21 + t2 + t4) * 2 + tvv * 4
        BB_k = reorder_fft_grid2(torch.fft.ifft2(BB_r, norm="forward"))
        return {"BB0_r": ((t1 + t21 + t2 + t4) * 2).real, "BB_r": BB_r.real, "B_r": tv.real, "BB_k": BB_k.real}


def bond_corr(BB_r_mean, B_r_mean):
    """BB_r - 4 * (autocorrelation of B_r) (stochastic_estimator.py:2643-2651)."""
    v_neg = torch.fft.ifftn(B_r_mean, norm="backward")
    v = torch.fft.fftn(B_r_mean, norm="forward")
    return BB_r_mean - 4 * torch.fft.ifftn(v_neg * v, norm="forward")


__all__ = ["BondBondEstimator", "bond_corr", "reorder_fft_grid2", "reorder_fft_grid2_inverse"]
