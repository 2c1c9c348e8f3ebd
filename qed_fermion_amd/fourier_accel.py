"""Fourier-accelerated momenta of the reference sampler: the adaptive mass estimate sigma_hat and
its two operators.

The reference (hmc_sampler_batch.py:268-285, 1133-1277, 3019-3033) estimates, per chain, the power
spectrum of the boson field's fluctuations over the run (``update_sigma_hat_cpu``), turns it every
``sigma_mini_batch_size`` sweeps into a normalised mass spectrum (``apply_sigma_hat_cpu``), and
defines the momentum draw p ~ N(0, M) with M = F^-1 diag(1/sigma_hat) F (``draw_momentum_fft``)
and the drift operator M^-1 p = F^-1 sigma_hat F p (``apply_m_inv``).  Its trajectory draws plain
momenta: the calls to ``draw_momentum_fft`` are commented out (2390, 2651) and ``apply_m_inv`` is
not called, so the estimate runs (when ``mass_mode`` != 0) without entering the dynamics.  The
same holds here: these are the reference's utilities with its semantics, and the trajectory keeps
p ~ N(0, m).

Spectra use the real-to-complex layout [bs, 2, Lx, Ly, Ltau // 2 + 1] of ``torch.fft.rfftn`` over
(x, y, tau).  The transforms run once per sweep on the sampler's device (rocFFT through torch.fft),
off the per-iteration hot path.
"""
import torch


class SigmaHatState(object):
    """Adaptive-mass state of one sampler (hmc_sampler_batch.py:271-285)."""

    def __init__(self, bs, Lx, Ly, Ltau, device, dtype, sigma_mini_batch_size=1000, anneal_times=10,
                 lmd=0.9, sig_min=0.8, sig_max=1.2, lambda_=0.05):
        shp = (bs, 2, Lx, Ly, Ltau // 2 + 1)
        self.dims = (Lx, Ly, Ltau)
        self.sigma_hat = torch.ones(shp, device=device, dtype=dtype)
        self.sigma_hat_cpu = torch.ones(shp, dtype=dtype)
        self.boson_mean_cpu = torch.zeros((bs, 2, Lx, Ly, Ltau), dtype=dtype)
        self.sigma_mini_batch_size = sigma_mini_batch_size
        self.sigma_hat_mini_batch_last_i = 0
        self.annealing_step = max(50, sigma_mini_batch_size)
        self.anneal_times = anneal_times
        # the real-to-complex half spectrum holds each mode once, except the first and the last
        # tau-frequency planes
        self.multiplier = torch.full_like(self.sigma_hat, 2)
        self.multiplier[..., torch.tensor([0, -1], dtype=torch.int64, device=self.sigma_hat.device)] = 1
        self.lmd = lmd
        self.sig_min = sig_min
        self.sig_max = sig_max
        self.lambda_ = lambda_


def update_sigma_hat(st, boson_cpu, i):
    """update_sigma_hat_cpu (hmc_sampler_batch.py:1202-1210): running mean of the field since the
    last annealing restart and running mean of |rfft(boson - mean)|^2 (orthonormal transform).  A
    restart happens every ``annealing_step`` sweeps, ``anneal_times`` times."""
    if i % st.annealing_step == 0 and i > 0 and st.anneal_times > 0:
        st.sigma_hat_mini_batch_last_i = i
        st.anneal_times -= 1
    cnt = i - st.sigma_hat_mini_batch_last_i
    st.boson_mean_cpu = (st.boson_mean_cpu * cnt + boson_cpu) / (cnt + 1)
    delta = torch.fft.rfftn(boson_cpu - st.boson_mean_cpu, st.dims, norm="ortho").abs() ** 2
    st.sigma_hat_cpu = (st.sigma_hat_cpu * cnt + delta) / (cnt + 1)


def stabilize_sigma_hat(sigma_hat, method="shrink", **kwargs):
    """stabilize_sigma_hat (hmc_sampler_batch.py:1212-1245); returns the new tensor."""
    if method == "shrink":
        lmd = kwargs.get("lmd", 0.8)
        return (1 - lmd) * sigma_hat + lmd * torch.ones_like(sigma_hat)
    if method == "pow":
        return sigma_hat.pow(kwargs.get("beta", 0.5))
    if method == "clamp":
        return sigma_hat.clamp(kwargs.get("sgm_min", 0.5), kwargs.get("sgm_max", 2.0))
    if method == "clamp_shrink":
        sgm_min, sgm_max = kwargs.get("sgm_min", 0.5), kwargs.get("sgm_max", 2.0)
        red = (2, 3, 4)
        max_val = sigma_hat.amax(dim=red, keepdim=True)
        min_val = sigma_hat.amin(dim=red, keepdim=True)
        return torch.where(sigma_hat > 1, 1 + (sigma_hat - 1) * ((sgm_max - 1) / (max_val - 1)),
                           1 - (1 - sigma_hat) * ((sgm_min - 1) / (min_val - 1)))
    if method == "log_inter":
        lmd = kwargs.get("lmd", 0.8)
        return (1 - lmd) * torch.log1p(sigma_hat) + lmd * torch.ones_like(sigma_hat)
    return sigma_hat


def apply_sigma_hat(st, i, mass_mode=0):
    """apply_sigma_hat_cpu (hmc_sampler_batch.py:1133-1200): every ``sigma_mini_batch_size`` sweeps
    the estimate becomes the mass spectrum: clamped at the mean over (chain, component) of the 0.99
    quantile, Box-Cox transformed with lambda 0.05, scaled to [sig_min, sig_max], shifted to mean 1
    and shrunk towards 1 by lmd; inverted for mass_mode -1.  Returns True when it was applied."""
    if not (i % st.sigma_mini_batch_size == 0 and i > 0):
        return False
    s = st.sigma_hat_cpu.to(st.sigma_hat.device)
    flat = s.view(s.size(0), s.size(1), -1)
    q = torch.quantile(flat, 0.99, dim=-1, keepdim=True).view(s.size(0), s.size(1), 1, 1, 1)
    s = stabilize_sigma_hat(s, "clamp", sgm_min=0, sgm_max=q.mean())
    st.lambda_ = 0.05   # fixed here by the reference, whatever the environment says
    s = (s ** st.lambda_ - 1) / st.lambda_
    red = (2, 3, 4)
    lo = s.amin(dim=red, keepdim=True)
    hi = s.amax(dim=red, keepdim=True)
    s = (s - lo) / (hi - lo).clamp(min=1e-8) * (st.sig_max - st.sig_min) + st.sig_min
    s = s - s.mean(dim=red, keepdim=True) + 1
    s = stabilize_sigma_hat(s, "shrink", lmd=st.lmd)
    if mass_mode == -1:
        s = s ** (-1)
    st.sigma_hat = s
    return True


def draw_momentum_fft(st, generator=None):
    """draw_momentum_fft (hmc_sampler_batch.py:1254-1265): p = irfft(z / sqrt(sigma_hat * multiplier))
    with z of independent unit normal real and imaginary parts, orthonormal transform, so that
    p ~ N(0, M), M = F^-1 diag(1 / sigma_hat) F."""
    shp = st.sigma_hat.shape
    dev, dt = st.sigma_hat.device, st.sigma_hat.dtype
    x = torch.randn(shp, device=dev, dtype=dt, generator=generator)
    y = torch.randn(shp, device=dev, dtype=dt, generator=generator)
    z = (x + 1j * y) / torch.sqrt(st.sigma_hat * st.multiplier)
    return torch.fft.irfftn(z, st.dims, norm="ortho")


def apply_m_inv(st, p):
    """apply_m_inv (hmc_sampler_batch.py:1267-1277): M^-1 p = F^-1 sigma_hat F p for p
    [bs, 2, Lx, Ly, Ltau] (the transform pair of torch.fft's default normalisation)."""
    return torch.fft.irfftn(torch.fft.rfftn(p, st.dims) * st.sigma_hat, st.dims)
