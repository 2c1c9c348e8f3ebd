"""The adaptive-mass utilities (qed_fermion_amd/fourier_accel.py, hmc_sampler_batch.py:1133-1277) on
CPU tensors, pinned by properties the formulas must satisfy rather than by restating them:
the transform pair (identity, a single Fourier mode's closed-form projector), the consistency of
the draw with M^-1 (equipartition), Parseval for the fluctuation spectrum, the annealing restart,
and the normalisation of the mass spectrum (mean 1, range, monotonicity)."""
import math

import numpy as np
import pytest
import torch

from qed_fermion_amd import fourier_accel as FA


def _state(bs=2, L=4, Lt=8, dtype=torch.float64, **kw):
    return FA.SigmaHatState(bs, L, L, Lt, torch.device("cpu"), dtype, **kw)


def test_m_inv_identity_and_scale():
    st = _state()
    p = torch.randn(2, 2, 4, 4, 8, dtype=torch.float64)
    assert torch.allclose(FA.apply_m_inv(st, p), p, atol=1e-13)
    st.sigma_hat = torch.full_like(st.sigma_hat, 2.5)
    assert torch.allclose(FA.apply_m_inv(st, p), 2.5 * p, atol=1e-12)


@pytest.mark.parametrize("k", [(1, 2, 3), (0, 0, 1), (3, 1, 2)])
def test_m_inv_single_mode_is_the_cosine_projector(k):
    """sigma_hat = 1 on one half-spectrum bin (0 < k_tau < Ltau/2) and 0 elsewhere: M^-1 projects on
    span{cos(k.r), sin(k.r)}, i.e. (M^-1 p)(r) = (2/N) sum_r' p(r') cos(k.(r - r'))."""
    L, Lt = 4, 8
    st = _state(bs=1, L=L, Lt=Lt)
    st.sigma_hat = torch.zeros_like(st.sigma_hat)
    st.sigma_hat[..., k[0], k[1], k[2]] = 1.0
    rng = np.random.default_rng(1)
    p = rng.standard_normal((1, 2, L, L, Lt))
    got = FA.apply_m_inv(st, torch.from_numpy(p)).numpy()
    x, y, t = np.meshgrid(np.arange(L), np.arange(L), np.arange(Lt), indexing="ij")
    ph = 2 * np.pi * (k[0] * x / L + k[1] * y / L + k[2] * t / Lt)
    N = L * L * Lt
    c, s = np.cos(ph), np.sin(ph)
    for comp in range(2):
        q = p[0, comp]
        want = (2.0 / N) * (c * np.sum(q * c) + s * np.sum(q * s))
        assert np.allclose(got[0, comp], want, atol=1e-12)


def test_draw_is_consistent_with_m_inv():
    """p ~ N(0, M) and M^-1 from the same sigma_hat: E[p^T M^-1 p] = N per (chain, component)
    (chi-square with N degrees of freedom), for a non-uniform spectrum that is symmetric under
    k -> -k as every |rfft|^2 of a real field is.  A draw scaled by sqrt(sigma_hat) instead of
    1/sqrt(sigma_hat), or without the half-spectrum multiplier, misses by far more than the bound."""
    torch.manual_seed(0)
    bs, L, Lt = 64, 4, 8
    N = L * L * Lt
    st = _state(bs=bs, L=L, Lt=Lt)
    w = torch.randn(1, 2, L, L, Lt, dtype=torch.float64)
    st.sigma_hat = (torch.fft.rfftn(w, (L, L, Lt), norm="ortho").abs() ** 2 + 0.3).expand(bs, -1, -1, -1, -1).contiguous()
    e = []
    for _ in range(8):
        p = FA.draw_momentum_fft(st)
        e.append((p * FA.apply_m_inv(st, p)).sum(dim=(2, 3, 4)).reshape(-1))
    e = torch.cat(e)
    n = e.numel()
    assert abs(e.mean().item() - N) < 5 * math.sqrt(2 * N / n), e.mean().item()
    # unit spectrum: unit variance per site
    st.sigma_hat = torch.ones_like(st.sigma_hat)
    v = torch.stack([FA.draw_momentum_fft(st) for _ in range(4)])
    assert abs(v.pow(2).mean().item() - 1.0) < 5 * math.sqrt(2.0 / v.numel())


def test_update_sigma_hat_running_means_and_parseval():
    """boson_mean_cpu is the arithmetic mean of the fields since the last restart, and the weighted
    half spectrum of sigma_hat_cpu sums (Parseval) to the running mean of |b_i - mean_i|^2."""
    torch.manual_seed(1)
    L, Lt = 4, 8
    st = _state(bs=2, L=L, Lt=Lt)
    fields, sq = [], []
    for i in range(7):
        b = torch.randn(2, 2, L, L, Lt, dtype=torch.float64)
        fields.append(b)
        FA.update_sigma_hat(st, b, i)
        m = torch.stack(fields).mean(0)
        assert torch.allclose(st.boson_mean_cpu, m, atol=1e-13)
        sq.append(((b - m) ** 2).sum(dim=(2, 3, 4)))
        parseval = (st.sigma_hat_cpu * st.multiplier).sum(dim=(2, 3, 4))
        assert torch.allclose(parseval, torch.stack(sq).mean(0), rtol=1e-12, atol=1e-12)


def test_update_sigma_hat_annealing_restart():
    L, Lt = 4, 8
    st = _state(bs=1, L=L, Lt=Lt, sigma_mini_batch_size=10, anneal_times=2)   # annealing_step = 50
    assert st.annealing_step == 50
    for i in range(50):
        FA.update_sigma_hat(st, torch.randn(1, 2, L, L, Lt, dtype=torch.float64), i)
    b = torch.randn(1, 2, L, L, Lt, dtype=torch.float64)
    FA.update_sigma_hat(st, b, 50)
    assert st.anneal_times == 1 and st.sigma_hat_mini_batch_last_i == 50
    assert torch.equal(st.boson_mean_cpu, b)
    assert torch.count_nonzero(st.sigma_hat_cpu) == 0


@pytest.mark.parametrize("mass_mode", [0, 1, -1])
def test_apply_sigma_hat_normalisation(mass_mode):
    """After apply_sigma_hat the spectrum of each (chain, component) has mean 1, a range of
    (1 - lmd)(sig_max - sig_min), and is a non-decreasing function of the estimate; mass_mode -1
    gives its reciprocal.  Nothing happens off the mini-batch boundary."""
    torch.manual_seed(2)
    st = _state(bs=2, L=4, Lt=8, sigma_mini_batch_size=20, lmd=0.9, sig_min=0.8, sig_max=1.2)
    st.sigma_hat_cpu = torch.rand_like(st.sigma_hat_cpu) ** 3 * 10
    before = st.sigma_hat.clone()
    assert not FA.apply_sigma_hat(st, 19, mass_mode) and torch.equal(st.sigma_hat, before)
    assert not FA.apply_sigma_hat(st, 0, mass_mode)
    assert FA.apply_sigma_hat(st, 20, mass_mode)
    s = st.sigma_hat if mass_mode != -1 else 1.0 / st.sigma_hat
    red = (2, 3, 4)
    assert torch.allclose(s.mean(dim=red), torch.ones(2, 2, dtype=torch.float64), atol=1e-12)
    rng = (s.amax(dim=red) - s.amin(dim=red))
    assert torch.allclose(rng, torch.full_like(rng, 0.1 * 0.4), atol=1e-12)
    for c in range(2):
        for comp in range(2):
            order = torch.argsort(st.sigma_hat_cpu[c, comp].reshape(-1))
            out = s[c, comp].reshape(-1)[order]
            assert bool((out[1:] >= out[:-1] - 1e-12).all())


def test_stabilize_methods():
    s = torch.tensor([[[[[0.2, 1.0, 3.0]]]]], dtype=torch.float64)
    assert torch.allclose(FA.stabilize_sigma_hat(s, "shrink", lmd=0.75), 0.25 * s + 0.75)
    assert torch.allclose(FA.stabilize_sigma_hat(s, "clamp", sgm_min=0.5, sgm_max=2.0), s.clamp(0.5, 2.0))
    cs = FA.stabilize_sigma_hat(s, "clamp_shrink", sgm_min=0.5, sgm_max=2.0)
    assert abs(cs.min().item() - 0.5) < 1e-12 and abs(cs.max().item() - 2.0) < 1e-12 and abs(cs[..., 1].item() - 1.0) < 1e-12
    assert torch.allclose(FA.stabilize_sigma_hat(s, "pow", beta=0.5), s.sqrt())
