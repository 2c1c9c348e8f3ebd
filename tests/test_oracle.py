"""Pins the FP64 CPU oracle (oracle/qf_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake fails it: a wrong bond group,
phase site, sign or cosh/sinh weight breaks the closed form D_g = expm(a_g K_g)
or the dense product; a wrong antiperiodic sign breaks det M = det(I + prod B);
a wrong phase convention breaks gauge covariance; a dropped force term breaks
the finite differences; a mistyped band constant breaks the reference's own
preconditioner dump.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg

import dense_ref as D
from helpers import make_phi, make_vec, relerr
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _rng(s):
    return np.random.default_rng(s)


@pytest.mark.parametrize("L", [2, 4, 6])
def test_D_factor_is_exponential_of_hopping(L):
    """D_g = cosh(a) + sinh(a) K_g equals expm(a K_g) because K_g^2 = 1 on a perfect matching."""
    phi_s = _rng(L).uniform(-np.pi, np.pi, (L, L, 2))
    for g in (1, 2, 3, 4):
        K = D.K_mat(L, phi_s, g)
        assert np.allclose(K @ K, np.eye(L * L), atol=1e-14)
        assert np.allclose(K, K.conj().T)
        a = D.a_g(g)
        assert np.abs(D.D_mat(L, phi_s, g) - scipy.linalg.expm(a * K)).max() < 1e-14


@pytest.mark.parametrize("L", [2, 4, 6, 8])
def test_B_matches_dense_product_and_is_hermitian(L):
    phi_s = _rng(10 + L).uniform(-np.pi, np.pi, (L, L, 2))
    B = D.B_mat(L, phi_s)
    Bo = np.stack([O.apply_B(phi_s, e, L) for e in np.eye(L * L, dtype=np.complex128)], axis=1)
    assert np.abs(Bo - B).max() < 1e-14
    assert np.abs(Bo - Bo.conj().T).max() < 1e-14


def test_trotter_error_is_third_order():
    """B_tau is the symmetric split of exp(dtau T): the error scales as dtau^3.  A wrong a_g
    (e.g. a_1 = dtau/2) makes it first order."""
    L = 4
    phi_s = _rng(3).uniform(-np.pi, np.pi, (L, L, 2))
    T = sum(D.K_mat(L, phi_s, g) for g in (1, 2, 3, 4))
    errs = []
    for dtau in (0.1, 0.05):
        Bo = np.stack([O.apply_B(phi_s, e, L, dtau=dtau) for e in np.eye(L * L, dtype=np.complex128)], axis=1)
        errs.append(np.abs(Bo - scipy.linalg.expm(dtau * T)).max())
    assert 6.0 < errs[0] / errs[1] < 10.0, errs


@pytest.mark.parametrize("L,Lt", [(2, 4), (4, 6), (6, 10)])
def test_operator_matches_dense(L, Lt):
    phi = make_phi(L, Lt, "uniform", seed=L * Lt)[0]
    x = make_vec(L, Lt, seed=2)[0]
    M = D.M_mat(L, Lt, phi)
    assert relerr(O.apply_M(phi, x, L, Lt), M @ x) < 1e-14
    assert relerr(O.apply_Mdag(phi, x, L, Lt), M.conj().T @ x) < 1e-14
    assert relerr(O.apply_MdagM(phi, x, L, Lt), M.conj().T @ (M @ x)) < 1e-14


@pytest.mark.parametrize("L,Lt", [(2, 4), (4, 4), (4, 8)])
def test_det_M_block_identity(L, Lt):
    """det M = det(I + B_{Nt-1} ... B_0): the antiperiodic corner carries +B (get_detM,
    fermion_added/hmc_sampler2_batch_fermion.py:537-585)."""
    phi = make_phi(L, Lt, "gauss", seed=5)[0]
    cols = [O.apply_M(phi, e, L, Lt) for e in np.eye(L * L * Lt, dtype=np.complex128)]
    Mo = np.stack(cols, axis=1)
    prod = np.eye(L * L, dtype=np.complex128)
    for tau in range(Lt):
        prod = D.B_mat(L, phi[tau]) @ prod
    d1 = np.linalg.slogdet(Mo)
    d2 = np.linalg.slogdet(np.eye(L * L) + prod)
    assert abs(d1[1] - d2[1]) < 1e-10 and abs(d1[0] - d2[0]) < 1e-10


def test_gauge_covariance():
    """phi_mu(i) -> phi_mu(i) + theta_i - theta_{i+mu} gives M -> U M U^dag, U = diag(e^{i theta})."""
    L, Lt = 4, 6
    rng = _rng(9)
    phi = make_phi(L, Lt, "uniform", seed=9)[0]
    theta = rng.uniform(-np.pi, np.pi, (L, L))   # [y, x], tau independent
    phi2 = phi.copy()
    phi2[..., 0] += theta - np.roll(theta, -1, axis=1)
    phi2[..., 1] += theta - np.roll(theta, -1, axis=0)
    U = np.tile(np.exp(1j * theta).ravel(), Lt)
    x = make_vec(L, Lt, seed=3)[0]
    assert relerr(O.apply_MdagM(phi2, U * x, L, Lt), U * O.apply_MdagM(phi, x, L, Lt)) < 1e-14
    F1 = O.force(phi, x, L, Lt)
    F2 = O.force(phi2, U * x, L, Lt)
    assert relerr(F2, F1) < 1e-12


def test_operator_at_zero_field():
    """At phi = 0 every D_g pair block is [[cosh a, sinh a], [sinh a, cosh a]]."""
    L = 4
    v = np.zeros(L * L, dtype=np.complex128)
    v[0] = 1.0
    w = O.apply_B(np.zeros((L, L, 2)), v, L, nfac=1)   # D4 only: site (0,0) is even -> g4 partner (0,-1)
    a = 0.05
    assert abs(w[0] - np.cosh(a)) < 1e-15
    assert abs(w[(L - 1) * L] - np.sinh(a)) < 1e-15
    assert np.count_nonzero(np.abs(w) > 1e-15) == 2


@pytest.mark.parametrize("L,Lt,precon", [(4, 30, True), (6, 10, False), (4, 26, True)])
def test_pcg_matches_direct_solve(L, Lt, precon):
    phi = make_phi(L, Lt, "gauss", seed=7)[0]
    b = make_vec(L, Lt, seed=8)[0]
    tab = O.precon_table(Lt) if precon else None
    x, it, err = O.pcg(phi, b, L, Lt, table=tab, rtol=1e-12, max_iter=2000)
    M = D.M_mat(L, Lt, phi)
    xs = np.linalg.solve(M.conj().T @ M, b)
    assert relerr(x, xs) < 1e-9
    assert err < 1e-12
    assert relerr(O.apply_MdagM(phi, x, L, Lt), b) < 1.01e-12


def test_pcg_iteration_counter_semantics():
    """Early exit (hmc_sampler_batch.py:828-835): the counter counts iterations after which the
    chain is still active; fixed-iteration mode runs max_iter and freezes x once converged."""
    L, Lt = 4, 30
    phi = make_phi(L, Lt, "gauss", seed=3)[0]
    b = make_vec(L, Lt, seed=4)[0]
    tab = O.precon_table(Lt)
    x1, it1, e1 = O.pcg(phi, b, L, Lt, table=tab, rtol=1e-6, max_iter=500, early_exit=True)
    x2, it2, e2 = O.pcg(phi, b, L, Lt, table=tab, rtol=1e-6, max_iter=500, early_exit=False)
    assert it1 == it2 and e1 == e2 and np.array_equal(x1, x2)
    x3, it3, e3 = O.pcg(phi, b, L, Lt, table=tab, rtol=1e-6, max_iter=it1 + 1, early_exit=False)
    assert e3 == e1 and it3 == it1
    x4, it4, e4 = O.pcg(phi, b, L, Lt, table=tab, rtol=1e-6, max_iter=it1, early_exit=False)
    assert e4 > 1e-6 and it4 == it1


def test_preconditioner_is_symmetric_band():
    Lt, Vs = 40, 4
    tab = O.precon_table(Lt)
    P = D.P_dense(tab, Vs)
    assert np.abs(P - P.T).max() == 0.0
    r = make_vec(2, Lt, seed=1)[0]
    assert relerr(O.apply_P(tab, r, Vs, Lt), P @ r) < 1e-15
    with pytest.raises(ValueError):
        O.precon_table(25)


def test_preconditioner_against_reference_dump():
    """precon_manual.py's edge/tail/corner constants are read off the ichol preconditioner dumped
    in preconditioners_orig/inspect_precon.md (L=8, Ltau=80): they must match it to float32
    rounding, and every band entry must agree with the dump within the reference's own
    manual-vs-computed tolerance (precon_manual.py debug branch: rtol=1e-2, atol=3e-3)."""
    d = json.load(open(os.path.join(HERE, "golden", "precon_dump_L8_Lt80.json")))
    Lt = d["Ltau"]
    tab = O.precon_table(Lt)
    n_edge = {0: 10, 1: 12, 2: 8, 3: 8, 4: 9, 5: 11, 6: 7}
    n_tail = {0: 9, 1: 8, 2: 7, 3: 6, 4: 8, 5: 10, 6: 7}
    for r in range(0, 7):
        dump = np.array(d["rank"][str(r)])
        man = tab[:, 6 + r]
        rows = np.arange(Lt)
        band = (rows < Lt - r) & (dump != -100)
        assert np.all(np.abs(man[band] - dump[band]) <= 3e-3 + 1e-2 * np.abs(dump[band])), r
        start = 1 if r == 6 else 0
        n = Lt - r - (1 if r == 6 else 0)
        for k in list(range(n_edge[r])) + list(range(n - n_tail[r], n)):
            assert abs(man[start + k] - dump[start + k]) < 1e-7, (r, k)
    for r in range(1, 6):
        assert abs(tab[0, 6 - r] - d["rank"][str(-r)][0]) < 1e-7


def _S_f(phi, psi, L, Lt):
    x, _, _ = O.pcg(phi, psi, L, Lt, rtol=1e-14, max_iter=5000)
    return float(np.real(np.vdot(psi, x)))


@pytest.mark.parametrize("L,Lt", [(2, 4), (4, 6)])
def test_force_against_dense_finite_differences(L, Lt):
    phi = make_phi(L, Lt, "uniform", seed=L + 17)[0]
    psi = make_vec(L, Lt, seed=5)[0]
    M = D.M_mat(L, Lt, phi)
    xi = np.linalg.solve(M.conj().T @ M, psi)
    F = O.force(phi, xi, L, Lt).ravel()
    rng = _rng(6)
    eps = 1e-5
    for k in rng.choice(phi.size, 12, replace=False):
        pp, pm = phi.copy().ravel(), phi.copy().ravel()
        pp[k] += eps
        pm[k] -= eps
        fd = -(D.S_f(L, Lt, pp.reshape(phi.shape), psi) - D.S_f(L, Lt, pm.reshape(phi.shape), psi)) / (2 * eps)
        assert abs(fd - F[k]) < 1e-6 * max(1.0, abs(F[k])), (k, fd, F[k])


def test_force_against_autograd():
    """The method of unit_test/test_force_f.py:31-45: autograd of S_f = psi^dag (M^dag M)^{-1} psi
    through a dense complex128 M built from the bond phases."""
    torch = pytest.importorskip("torch")
    L, Lt = 4, 4
    phi = make_phi(L, Lt, "uniform", seed=23)[0]
    psi = make_vec(L, Lt, seed=24)[0]
    lists = [(i.copy(), j.copy(), mu) for i, j, mu in D.bond_lists(L)]
    ph = torch.tensor(phi, dtype=torch.float64, requires_grad=True)
    Vs = L * L
    Ms = torch.eye(Vs * Lt, dtype=torch.complex128)
    blocks = []
    for tau in range(Lt):
        Dm = {}
        for g in (1, 2, 3, 4):
            i, j, mu = lists[g - 1]
            a = D.a_g(g)
            e = torch.exp(1j * ph[tau].reshape(-1, 2)[torch.as_tensor(i), mu])
            K = torch.zeros(Vs, Vs, dtype=torch.complex128)
            K = K.index_put((torch.as_tensor(i), torch.as_tensor(j)), e)
            K = K.index_put((torch.as_tensor(j), torch.as_tensor(i)), e.conj())
            Dm[g] = np.cosh(a) * torch.eye(Vs, dtype=torch.complex128) + np.sinh(a) * K
        B = Dm[1]
        for g in (2, 3, 4):
            B = Dm[g] @ B @ Dm[g]
        blocks.append(B)
    rows = []
    for tau in range(Lt):
        row = [torch.zeros(Vs, Vs, dtype=torch.complex128) for _ in range(Lt)]
        row[tau] = torch.eye(Vs, dtype=torch.complex128)
        if tau > 0:
            row[tau - 1] = -blocks[tau - 1]
        else:
            row[Lt - 1] = blocks[Lt - 1]
        rows.append(torch.cat(row, dim=1))
    Ms = torch.cat(rows, dim=0)
    A = Ms.conj().T @ Ms
    ps = torch.tensor(psi)
    S = torch.real(torch.vdot(ps, torch.linalg.solve(A, ps)))
    S.backward()
    xi = np.linalg.solve(A.detach().numpy(), psi)
    F = O.force(phi, xi, L, Lt)
    assert relerr(F, -ph.grad.numpy()) < 1e-10


def test_boson_actions_closed_forms_and_forces():
    L, Lt = 2, 2
    cJ, cK = O.couplings(1.0, 1.0, 0.1)
    pf = make_phi(L, Lt, "piflux")[0]
    assert abs(O.action_plaq(pf, L, Lt, cK) - (-cK * L * L * Lt)) < 1e-14   # -8K on 2x2x2 (test_action_plaq.py:49-52)
    assert O.action_tau(pf, L, Lt, cJ) == 0.0
    L, Lt = 4, 6
    phi = make_phi(L, Lt, "gauss", seed=2)[0]
    for compact in (False, True):
        Fp = O.force_b_plaq(phi, L, Lt, cK).ravel()
        Ft = O.force_b_tau(phi, L, Lt, cJ, compact).ravel()
        eps = 1e-6
        for k in _rng(1).choice(phi.size, 10, replace=False):
            pp, pm = phi.copy().ravel(), phi.copy().ravel()
            pp[k] += eps
            pm[k] -= eps
            dSp = (O.action_plaq(pp.reshape(phi.shape), L, Lt, cK) - O.action_plaq(pm.reshape(phi.shape), L, Lt, cK)) / (2 * eps)
            dSt = (O.action_tau(pp.reshape(phi.shape), L, Lt, cJ, compact) -
                   O.action_tau(pm.reshape(phi.shape), L, Lt, cJ, compact)) / (2 * eps)
            assert abs(-dSp - Fp[k]) < 1e-6
            assert abs(-dSt - Ft[k]) < 1e-6 * max(1, abs(Ft[k]))


def test_trajectory_energy_error_second_order():
    """Leapfrog at fixed trajectory length: Delta H = O(dt^2) (ratio ~4 when dt halves)."""
    L, Lt = 4, 30
    phi = make_phi(L, Lt, "gauss", seed=4)[0]
    p = _rng(5).standard_normal(phi.shape)
    R = make_vec(L, Lt, seed=6)[0]
    tab = O.precon_table(Lt)
    kw = dict(table=tab, rtol=1e-13, max_iter=3000)
    dH = []
    for dt, n in ((0.02, 4), (0.01, 8), (0.005, 16)):
        _, _, a, b, err = O.trajectory(phi, p, R, L, Lt, dt=dt, n_leap=n, **kw)
        assert err < 1e-12
        dH.append(abs(b - a))
    assert 3.0 < dH[0] / dH[1] < 5.5 and 3.0 < dH[1] / dH[2] < 5.5, dH


def test_trajectory_reversible():
    """Integrating back with -p returns to the start (the fermion force uses the same psi)."""
    L, Lt = 4, 30
    phi = make_phi(L, Lt, "gauss", seed=7)[0]
    p = _rng(8).standard_normal(phi.shape)
    R = make_vec(L, Lt, seed=9)[0]
    tab = O.precon_table(Lt)
    kw = dict(table=tab, rtol=1e-14, max_iter=4000, dt=0.02, n_leap=3)
    psi = O.apply_Mdag(phi, R, L, Lt)
    x1, p1, _, _, _ = O.trajectory(phi, p, R, L, Lt, **kw)
    # psi is fixed along the trajectory; express it through R' = M(x1)^{-dag} psi at the end point
    Mx1 = np.stack([O.apply_Mdag(x1, e, L, Lt) for e in np.eye(L * L * Lt, dtype=np.complex128)], axis=1)
    R1 = np.linalg.solve(Mx1, psi)
    x2, p2, _, _, _ = O.trajectory(x1, -p1, R1, L, Lt, **kw)
    assert relerr(x2, phi) < 1e-10
    assert relerr(-p2, p) < 1e-10
