"""Pins of the bond-bond estimator oracle (oracle/observables.py), CPU only.

* The three-FFT correlation of the reference equals the direct sum
  (1/V) sum_r a(r - Delta) b(r) (brute force, a non-square grid).
* Gauge invariance: under a static gauge transformation
  phi_mu(i) -> phi_mu(i) + theta_i - theta_{i+mu}, M -> U M U^dag (checked on the
  dense M), so (eta, G_eta) -> (U eta, U G_eta) and every estimator must be
  unchanged sample by sample.  A wrong neighbour or phase sign in any term
  breaks this.
* Exact noise average: the pair part of T1/T21/T2/T4 averaged over noise
  vectors (phase * unit vector, phases {1, i, -1, -i}, which reproduces
  E[eta* eta] = 1, E[eta eta] = 0 exactly for one quadratic form per noise
  vector) equals the exchange contraction of the dense M^{-1} written as site
  loops: for T1, -(1/(Ltau V)) sum_{tau,r} C(r+x, r-D) C(r-D+x, r) U(r-D) U(r)
  with C = M^{-1}.  This fixes the conjugation, the sign, the normalisation and
  the direction of Delta.
"""
import numpy as np
import pytest

import dense_ref as D
from helpers import make_phi, relerr
from oracle import observables as OB


def _boson(phi):
    """oracle layout phi [Lt, L, L, 2] (tau, y, x, mu) -> reference boson [1, 2, Lx, Ly, Lt]."""
    return np.transpose(phi, (3, 2, 1, 0))[None]


def _z2(rng, nrv, N):
    return ((rng.integers(0, 2, (nrv, N)) * 2 - 1) + 1j * (rng.integers(0, 2, (nrv, N)) * 2 - 1)) / np.sqrt(2)


def test_fft_correlation_is_the_direct_sum():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((3, 2, 4, 6)) + 1j * rng.standard_normal((3, 2, 4, 6))
    b = rng.standard_normal((3, 2, 4, 6)) + 1j * rng.standard_normal((3, 2, 4, 6))
    assert relerr(OB.corr_mean(a, b), OB.corr_direct(a, b)) < 1e-13


def test_pairs_are_torch_combinations_order():
    import torch
    want = torch.combinations(torch.arange(7), r=2).numpy()
    assert np.array_equal(OB.pairs(7), want)


def _setup(L=4, Lt=6, nrv=5, seed=3):
    rng = np.random.default_rng(seed)
    phi = make_phi(L, Lt, "uniform", seed=seed)[0]
    M = D.M_mat(L, Lt, phi)
    eta = _z2(rng, nrv, L * L * Lt)
    G = np.linalg.solve(M, eta.T).T
    return rng, phi, M, eta, G


def test_estimators_are_gauge_invariant():
    L, Lt = 4, 6
    rng, phi, M, eta, G = _setup(L, Lt)
    theta = rng.uniform(-np.pi, np.pi, (L, L))
    phi2 = phi.copy()
    phi2[..., 0] += theta - np.roll(theta, -1, axis=1)
    phi2[..., 1] += theta - np.roll(theta, -1, axis=0)
    U = np.tile(np.exp(1j * theta).ravel(), Lt)
    M2 = D.M_mat(L, Lt, phi2)
    assert relerr(M2, U[:, None] * M * np.conj(U)[None, :]) < 1e-13      # the transformation law used below
    eta2, G2 = eta * U, G * U
    b1, b2 = _boson(phi), _boson(phi2)
    for name in ("T1", "T21", "T2", "T4"):
        t1 = OB.T(name, eta, G, b1, L, Lt)
        t2 = OB.T(name, eta2, G2, b2, L, Lt)
        assert relerr(t2, t1) < 1e-12, name
    v1, v2 = OB.Tvv(eta, G, b1, L, Lt), OB.Tvv(eta2, G2, b2, L, Lt)
    assert relerr(v2[0], v1[0]) < 1e-12 and relerr(v2[1], v1[1]) < 1e-12


def test_gauge_invariance_catches_a_wrong_neighbour(monkeypatch):
    # a planted mistake (T4's eta taken at r instead of r + x) must fail the invariance check
    L, Lt = 4, 6
    rng, phi, M, eta, G = _setup(L, Lt)
    theta = rng.uniform(-np.pi, np.pi, (L, L))
    phi2 = phi.copy()
    phi2[..., 0] += theta - np.roll(theta, -1, axis=1)
    phi2[..., 1] += theta - np.roll(theta, -1, axis=0)
    U = np.tile(np.exp(1j * theta).ravel(), Lt)
    orig = OB._ab

    def planted(name, e, G_, b, s, sp):
        a, bb = orig(name, e, G_, b, s, sp)
        if name == "T4":
            a = OB._roll(e[s], 0) * OB._roll(G_[sp], 0) * np.exp(-1j * b[0, 0])
        return a, bb

    monkeypatch.setattr(OB, "_ab", planted)
    t1 = OB.pair_part("T4", eta, G, _boson(phi), L, Lt)
    t2 = OB.pair_part("T4", eta * U, G * U, _boson(phi2), L, Lt)
    assert relerr(t2, t1) > 1e-3


def _exchange(name, Minv, phi, L, Lt):
    """Noise expectation of the pair part, from C = M^{-1} by site loops."""
    V = L * L
    Ux = np.exp(1j * phi[..., 0])          # [Lt, y, x]
    idx = lambda t, y, x: t * V + (y % L) * L + (x % L)
    C = lambda j, i: Minv[j, i]
    out = np.zeros((L, L), dtype=np.complex128)
    for dy in range(L):
        for dx in range(L):
            acc = 0.0
            for t in range(Lt):
                for y in range(L):
                    for x in range(L):
                        ry, rx = y - dy, x - dx                   # rho = r - Delta
                        r, rp = idx(t, y, x), idx(t, y, x + 1)
                        q, qp = idx(t, ry, rx), idx(t, ry, rx + 1)
                        Ur, Uq = Ux[t, y, x], Ux[t, ry % L, rx % L]
                        if name == "T1":
                            acc += C(rp, q) * C(qp, r) * Uq * Ur
                        elif name == "T4":
                            acc += C(r, qp) * C(q, rp) * np.conj(Uq) * np.conj(Ur)
                        elif name == "T21":
                            acc += 0.5 * C(rp, qp) * C(q, r) * np.conj(Uq) * Ur
                        elif name == "T2":
                            acc += 0.5 * C(r, q) * C(qp, rp) * Uq * np.conj(Ur)
            out[dy, dx] = -acc / (Lt * V)
    return out


@pytest.mark.parametrize("name", ["T1", "T21", "T2", "T4"])
def test_pair_part_noise_average_is_exchange_contraction(name):
    L, Lt = 4, 4
    N = L * L * Lt
    phi = make_phi(L, Lt, "uniform", seed=11)[0]
    M = D.M_mat(L, Lt, phi)
    Minv = np.linalg.inv(M)
    phases = np.array([1, 1j, -1, -1j])
    basis = (phases[:, None, None] * np.eye(N)[None]).reshape(4 * N, N)   # [4N, N]
    eta = np.concatenate([basis, basis])                                 # set 1 then set 2
    G = (Minv @ eta.T).T
    n = 4 * N
    s = np.repeat(np.arange(n), n)
    sp = n + np.tile(np.arange(n), n)                                    # every (set-1, set-2) pair, s < s'
    est = OB.pair_part(name, eta, G, _boson(phi), L, Lt, idx=np.stack([s, sp], axis=1))
    # E over two independent noise vectors = (1/16) sum over both sets = N^2 * mean over pairs
    assert relerr(est * N * N, _exchange(name, Minv, phi, L, Lt)) < 1e-10


def test_tvv_against_loop_over_noise_vectors():
    L, Lt = 4, 6
    rng, phi, M, eta, G = _setup(L, Lt, nrv=4)
    b = _boson(phi)
    tvv, tv = OB.Tvv(eta, G, b, L, Lt)
    v = OB.v_field(eta, G, b, L, Lt)
    want = np.mean([OB.corr_direct(v[s:s + 1], v[s:s + 1]) for s in range(4)], axis=0)
    assert relerr(tvv, want) < 1e-12
    assert relerr(tv, v.mean(axis=(0, 1))) < 1e-14


def test_bond_bond_combination_and_k_grid():
    L, Lt = 4, 6
    rng, phi, M, eta, G = _setup(L, Lt)
    out = OB.bond_bond(eta, G, _boson(phi), L, Lt)
    t = sum(OB.T(n, eta, G, _boson(phi), L, Lt) for n in ("T1", "T21", "T2", "T4"))
    tvv, _ = OB.Tvv(eta, G, _boson(phi), L, Lt)
    assert np.allclose(out["BB_r"], (2 * t + 4 * tvv).real)
    # BB_k at the grid centre (index L/2, L/2) is the k = 0 mode: sum over Delta of BB_r
    assert np.isclose(out["BB_k"][L // 2, L // 2], (2 * t + 4 * tvv).sum().real)


def test_tv_noise_average_is_the_exact_bond():
    """Tv over the exact noise basis (every unit vector with each of the phases 1, i, -1, -i:
    the sample mean of eta*_i eta_k is delta_ik / N) equals the dense M^{-1} bond of
    dense_ref.bond_x, the fermion observable of the sampler's statistics test; its diagonal
    blocks equal those of the dense inverse; a transposed M^{-1} (the plausible slip) differs,
    since M^{-1} is not Hermitian."""
    L, Lt = 4, 4
    N = L * L * Lt
    phi = make_phi(L, Lt, "uniform", seed=12)[0]
    Minv = np.linalg.inv(D.M_mat(L, Lt, phi))
    phases = np.array([1, 1j, -1, -1j])
    eta = (phases[:, None, None] * np.eye(N)[None]).reshape(4 * N, N)
    G = (Minv @ eta.T).T
    _, tv = OB.Tvv(eta, G, _boson(phi), L, Lt)
    want = D.bond_x(L, Lt, phi)
    assert relerr(tv * N, want) < 1e-10
    G = D.equal_time_G(L, Lt, phi)
    for tau in range(Lt):
        blk = Minv[tau * L * L:(tau + 1) * L * L, tau * L * L:(tau + 1) * L * L]
        assert relerr(G[tau], blk) < 1e-12
    tt, yy, xx = np.meshgrid(np.arange(Lt), np.arange(L), np.arange(L), indexing="ij")
    r, rx = tt * L * L + yy * L + xx, tt * L * L + yy * L + (xx + 1) % L
    u = np.exp(1j * phi[..., 0])
    slip = (Minv[r, rx] * u + Minv[rx, r] * np.conj(u)).mean(axis=0)
    assert relerr(slip, want) > 1e-2
