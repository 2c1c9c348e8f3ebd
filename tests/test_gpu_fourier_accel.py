"""The sampler's Fourier-accelerated momenta (draw_momentum_fft, apply_m_inv;
hmc_sampler_batch.py:1254-1277) on the device, pinned by the mathematics rather than by the
same code on the CPU: M^-1 on a single Fourier mode is the closed-form cosine projector, the
draw is consistent with M^-1 (E[p^T M^-1 p] = N per chain and component), a unit spectrum
gives unit variance, and M^-1 preserves the L2 norm weighted by the spectrum (Parseval)."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _sampler(L, Lt, bs):
    from qed_fermion_amd.hmc_sampler import HmcSampler
    return HmcSampler(Lx=L, Ltau=Lt, bs=bs, precision="fp64", device="cuda")


@pytest.mark.parametrize("k", [(1, 2, 3), (0, 0, 1), (3, 1, 19)])
def test_m_inv_single_mode_on_device(dev, k):
    """sigma_hat = 1 on one bin with 0 < k_tau < Ltau/2, 0 elsewhere: (M^-1 p)(r) =
    (2/N) sum_r' p(r') cos(k.(r - r'))."""
    L, Lt = 4, 40
    hmc = _sampler(L, Lt, 1)
    s = torch.zeros_like(hmc.sigma_hat)
    s[..., k[0], k[1], k[2]] = 1.0
    hmc.sigma_hat = s
    p = np.random.default_rng(5).standard_normal((1, 2, L, L, Lt))
    got = hmc.apply_m_inv(torch.from_numpy(p).to(dev))
    assert got.is_cuda
    got = got.cpu().numpy()
    x, y, t = np.meshgrid(np.arange(L), np.arange(L), np.arange(Lt), indexing="ij")
    ph = 2 * np.pi * (k[0] * x / L + k[1] * y / L + k[2] * t / Lt)
    c, sn = np.cos(ph), np.sin(ph)
    N = L * L * Lt
    for comp in range(2):
        q = p[0, comp]
        want = (2.0 / N) * (c * np.sum(q * c) + sn * np.sum(q * sn))
        assert np.abs(got[0, comp] - want).max() < 1e-12


def test_draw_consistent_with_m_inv_on_device(dev):
    """p ~ N(0, M) from the device draw: p^T M^-1 p is chi-square with N degrees of freedom per
    (chain, component) for a non-uniform spectrum symmetric under k -> -k."""
    L, Lt, bs = 4, 40, 64
    N = L * L * Lt
    hmc = _sampler(L, Lt, bs)
    torch.manual_seed(11)
    w = torch.randn(1, 2, L, L, Lt, dtype=torch.float64, device=dev)
    hmc.sigma_hat = (torch.fft.rfftn(w, (L, L, Lt), norm="ortho").abs() ** 2 + 0.3).expand(bs, -1, -1, -1, -1).contiguous()
    e = []
    for _ in range(4):
        p = hmc.draw_momentum_fft()
        assert p.is_cuda and p.shape == (bs, 2, L, L, Lt)
        e.append((p * hmc.apply_m_inv(p)).sum(dim=(2, 3, 4)).reshape(-1))
    e = torch.cat(e)
    assert abs(e.mean().item() - N) < 5 * math.sqrt(2 * N / e.numel()), e.mean().item()
    hmc.sigma_hat = torch.ones_like(hmc.sigma_hat)
    v = torch.stack([hmc.draw_momentum_fft() for _ in range(2)])
    assert abs(v.pow(2).mean().item() - 1.0) < 5 * math.sqrt(2.0 / v.numel())


def test_m_inv_parseval_on_device(dev):
    """<p, M^-1 p> = sum_k w_k sigma_k |p_hat(k)|^2 with the half-spectrum weights (1 for k_tau = 0
    and Ltau/2, 2 otherwise) and the orthonormal transform: M^-1 is diagonal in Fourier space."""
    L, Lt, bs = 6, 40, 3
    hmc = _sampler(L, Lt, bs)
    torch.manual_seed(12)
    w = torch.randn(1, 2, L, L, Lt, dtype=torch.float64, device=dev)   # a spectrum symmetric under k -> -k
    hmc.sigma_hat = (torch.fft.rfftn(w, (L, L, Lt), norm="ortho").abs() ** 2 + 0.5).expand(bs, -1, -1, -1, -1).contiguous()
    p = torch.randn(bs, 2, L, L, Lt, dtype=torch.float64, device=dev)
    lhs = (p * hmc.apply_m_inv(p)).sum(dim=(2, 3, 4))
    ph = torch.fft.rfftn(p, dim=(2, 3, 4), norm="ortho")
    wts = torch.full((Lt // 2 + 1,), 2.0, dtype=torch.float64, device=dev)
    wts[0] = wts[-1] = 1.0
    rhs = (wts * hmc.sigma_hat * ph.abs() ** 2).sum(dim=(2, 3, 4))
    assert torch.allclose(lhs, rhs, rtol=1e-12)
