"""Numpy restatement of the reference's bond-bond stochastic estimators.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
observable kernels.  The product package ``qed_fermion_amd`` never imports it.

Follows qed_fermion/stochastic_estimator.py:
  T1   1393-1445      T21  1447-1502      T2   1504-1559
  T4   1561-1610      Tvv  1612-1654      get_bond_bond_per_b2  2594-2641
  reorder_fft_grid2 253-256               bond_corr 2643-2651

Inputs (numpy): eta, G_eta complex [Nrv, Ltau*L*L] (index tau*L*L + y*L + x);
boson float [1, 2, L, L, Ltau] in the reference's storage layout
(boson[0, mu, x, y, tau]).  All pairs (s < s') are taken in
torch.combinations order in one batch; the reference's per-batch weighting
(end-start)/num_samples sums to the same mean.

Pinning (tests/test_observables_oracle.py):
  * corr_mean's three-FFT sequence against the direct sum
    (1/V) sum_r a(r - Delta) b(r) (brute force, no FFT);
  * unbiasedness: with many Z2 noise vectors at a tiny lattice each
    estimator's pair part approaches the Wick contraction of the dense
    M^{-1} (E[eta*_s(i) (M^{-1} eta_s)(j)] = M^{-1}(j, i), E[eta eta] = 0),
    within its measured standard error;
  * Tvv / Tv against a direct loop over s of the same v_s.
Which neighbour each factor takes (the rolls) and the phase signs are read
from the reference lines above; nothing else fixes them.
"""
import itertools

import numpy as np


def pairs(nrv):
    """(s, s') with s < s' in torch.combinations order."""
    return np.array(list(itertools.combinations(range(nrv), 2)), dtype=np.int64).reshape(-1, 2)


def _prep(eta, G_eta, boson, L, Lt):
    b = np.transpose(np.asarray(boson), (0, 1, 4, 3, 2))        # [1, 2, Ltau, Ly, Lx]
    e = np.conj(np.asarray(eta)).reshape(-1, Lt, L, L)
    G = np.asarray(G_eta).reshape(-1, Lt, L, L)
    return e, G, b


def _roll(t, k):
    """torch.roll(t, shifts=k, dims=-1): result[..., x] = t[..., (x - k) % L]."""
    return np.roll(t, k, axis=-1)


def corr_mean(a, b):
    """mean over (pair, tau) of ifftn(ifftn(a, 'backward') * fftn(b, 'forward'), 'forward')
    over the last two axes (stochastic_estimator.py:1428-1430)."""
    aF = np.fft.ifftn(a, axes=(-2, -1), norm="backward")
    bF = np.fft.fftn(b, axes=(-2, -1), norm="forward")
    r = np.fft.ifftn(aF * bF, axes=(-2, -1), norm="forward")
    return r.mean(axis=(0, 1))


def corr_direct(a, b):
    """(1/V) sum_r a(r - Delta) b(r), averaged over the leading two axes, by direct sums."""
    n0, n1, Ly, Lx = a.shape
    out = np.zeros((Ly, Lx), dtype=np.complex128)
    for dy in range(Ly):
        for dx in range(Lx):
            sh = np.roll(np.roll(a, dy, axis=-2), dx, axis=-1)   # sh[y, x] = a[y - dy, x - dx]
            out[dy, dx] = (sh * b).sum() / (Ly * Lx)
    return out / (n0 * n1)


def _ab(name, e, G, b, s, sp):
    U = np.exp(1j * b[0, 0])
    Ub = np.exp(-1j * b[0, 0])
    if name == "T1":
        a = _roll(e[s], 0) * _roll(G[sp], -1) * U
        bb = _roll(G[s], -1) * _roll(e[sp], 0) * U
    elif name == "T21":
        a = (_roll(e[s], -1) * _roll(G[sp], 0) + _roll(e[sp], -1) * _roll(G[s], 0)) * Ub / 2
        bb = (_roll(G[s], -1) * _roll(e[sp], 0) + _roll(G[sp], -1) * _roll(e[s], 0)) * U / 2
    elif name == "T2":
        a = (_roll(e[s], 0) * _roll(G[sp], -1) + _roll(e[sp], 0) * _roll(G[s], -1)) * U / 2
        bb = (_roll(G[s], 0) * _roll(e[sp], -1) + _roll(G[sp], 0) * _roll(e[s], -1)) * Ub / 2
    elif name == "T4":
        a = _roll(e[s], -1) * _roll(G[sp], 0) * Ub
        bb = _roll(G[s], 0) * _roll(e[sp], -1) * Ub
    else:
        raise ValueError(name)
    return a, bb


def pair_part(name, eta, G_eta, boson, L, Lt, idx=None):
    """-mean over pairs and tau of the correlation of a and b: the FFT part of T1/T21/T2/T4."""
    e, G, b = _prep(eta, G_eta, boson, L, Lt)
    idx = pairs(e.shape[0]) if idx is None else idx
    a, bb = _ab(name, e, G, b, idx[:, 0], idx[:, 1])
    return -corr_mean(a, bb)


def diag_part(name, eta, G_eta, boson, L, Lt):
    """(index, value) of the diagonal term each estimator adds to one Delta."""
    e, G, b = _prep(eta, G_eta, boson, L, Lt)
    U = np.exp(1j * b[0, 0])
    Ub = np.exp(-1j * b[0, 0])
    if name == "T1":      # 1434-1440
        d = _roll(G, -1) * _roll(e, 1) * U * np.exp(1j * _roll(b, 1)[0, 0])
        return (0, L - 1), d.mean()
    if name in ("T21", "T2"):   # 1496-1500, 1553-1557
        return (0, 0), (G * e).mean()
    if name == "T4":      # 1601-1608
        d = _roll(G, 0) * _roll(e, -2) * Ub * np.exp(-1j * _roll(b, -1)[0, 0])
        return (0, 1), d.mean()
    raise ValueError(name)


def T(name, eta, G_eta, boson, L, Lt):
    out = pair_part(name, eta, G_eta, boson, L, Lt)
    (iy, ix), d = diag_part(name, eta, G_eta, boson, L, Lt)
    out[iy, ix] += d
    return out


def v_field(eta, G_eta, boson, L, Lt):
    """v_s of Tvv (1631-1637): [Nrv, Ltau, Ly, Lx]."""
    e, G, b = _prep(eta, G_eta, boson, L, Lt)
    return (_roll(G, -1) * _roll(e, 0) * np.exp(1j * b[0, 0])
            + _roll(G, 0) * _roll(e, -1) * np.exp(-1j * b[0, 0]))


def Tvv(eta, G_eta, boson, L, Lt):
    """(Tvv, Tv): mean over (s, tau) of corr(v_s, v_s), and of v_s itself (1612-1654)."""
    v = v_field(eta, G_eta, boson, L, Lt)
    return corr_mean(v, v), v.mean(axis=(0, 1))


def reorder_fft_grid2(t):
    Ly, Lx = t.shape[-2:]
    return np.roll(t, (Ly // 2, Lx // 2), axis=(-2, -1))


def bond_bond(eta, G_eta, boson, L, Lt):
    """get_bond_bond_per_b2 (2594-2641): {'BB0_r', 'BB_r', 'B_r', 'BB_k'} as [Ly, Lx] reals."""
    t = {n: T(n, eta, G_eta, boson, L, Lt) for n in ("T1", "T21", "T2", "T4")}
    tvv, tv = Tvv(eta, G_eta, boson, L, Lt)
    s = t["T1"] + t["T21"] + t["T2"] + t["T4"]
    BB_r = s * 2 + tvv * 4
    BB_k = reorder_fft_grid2(np.fft.ifft2(BB_r, norm="forward"))
    return {"BB0_r": (s * 2).real, "BB_r": BB_r.real, "B_r": tv.real, "BB_k": BB_k.real}


def bond_corr(BB_r_mean, B_r_mean):
    """bond_corr (2643-2651): BB_r - 4 * (autocorrelation of B_r)."""
    vF_neg = np.fft.ifftn(B_r_mean, norm="backward")
    vF = np.fft.fftn(B_r_mean, norm="forward")
    v_bg = np.fft.ifftn(vF_neg * vF, norm="forward")
    return BB_r_mean - 4 * v_bg
