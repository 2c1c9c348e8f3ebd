"""Dense numpy brute force for tiny lattices: the independent construction the
oracle is pinned against.

Builds the checkerboard bond lists the way hmc_sampler_batch.py:937-969 does
(index arithmetic over even/odd sub-lattices), dense D_g and B_tau matrices by
explicit matrix products (get_M, hmc_sampler_batch.py:1552-1611) and the dense
block-bidiagonal M with the antiperiodic corner.  Nothing here shares code with
oracle/qf_oracle.c, which works matrix-free with pair updates.
"""
import functools

import numpy as np


@functools.lru_cache(maxsize=None)
def bond_lists(L):
    """[(i_list, j_list, mu)] for groups 1..4, sites linearised as y*L + x (cached per L; the
    arrays are read-only)."""
    xs = np.arange(0, L, 2)[None, :]
    ys = np.arange(0, L, 2)[:, None]
    xs2 = np.arange(1, L, 2)[None, :]
    ys2 = np.arange(1, L, 2)[:, None]
    even = np.concatenate([(xs + ys * L).ravel(), (xs2 + ys2 * L).ravel()])
    xe = np.concatenate([np.broadcast_to(xs, (len(ys), len(xs[0]))).ravel(),
                         np.broadcast_to(xs2, (len(ys2), len(xs2[0]))).ravel()])
    ye = np.concatenate([np.broadcast_to(ys, (len(ys), len(xs[0]))).ravel(),
                         np.broadcast_to(ys2, (len(ys2), len(xs2[0]))).ravel()])
    g1 = (even, (xe + 1) % L + ye * L, 0)
    g2 = (even, xe + ((ye + 1) % L) * L, 1)
    g3 = ((xe - 1) % L + ye * L, even, 0)
    g4 = (xe + ((ye - 1) % L) * L, even, 1)
    for g in (g1, g2, g3, g4):
        for a in g[:2]:
            a.setflags(write=False)
    return [g1, g2, g3, g4]


def K_mat(L, phi_s, g):
    """Hermitian hopping matrix of group g (1-based) on one slice; phi_s [L,L,2] (y,x,mu)."""
    i, j, mu = bond_lists(L)[g - 1]
    ph = phi_s.reshape(-1, 2)[i, mu]
    K = np.zeros((L * L, L * L), dtype=np.complex128)
    K[i, j] = np.exp(1j * ph)
    K[j, i] = np.exp(-1j * ph)
    return K


def a_g(g, dtau=0.1, t=1.0):
    return t * dtau if g == 1 else t * dtau / 2


def D_mat(L, phi_s, g, dtau=0.1, t=1.0):
    a = a_g(g, dtau, t)
    return np.cosh(a) * np.eye(L * L) + np.sinh(a) * K_mat(L, phi_s, g)


def B_mat(L, phi_s, dtau=0.1, t=1.0):
    D = {g: D_mat(L, phi_s, g, dtau, t) for g in (1, 2, 3, 4)}
    B = D[1]
    for g in (2, 3, 4):
        B = D[g] @ B @ D[g]
    return B


def M_mat(L, Lt, phi, dtau=0.1, t=1.0):
    """Dense M, phi [Lt, L, L, 2]."""
    Vs = L * L
    M = np.eye(Vs * Lt, dtype=np.complex128)
    for tau in range(Lt):
        B = B_mat(L, phi[tau], dtau, t)
        if tau < Lt - 1:
            M[Vs * (tau + 1):Vs * (tau + 2), Vs * tau:Vs * (tau + 1)] = -B
        else:
            M[:Vs, Vs * tau:] = B
    return M


def S_f(L, Lt, phi, psi, dtau=0.1, t=1.0):
    M = M_mat(L, Lt, phi, dtau, t)
    A = M.conj().T @ M
    return float(np.real(np.vdot(psi, np.linalg.solve(A, psi))))


def P_dense(table, Vs):
    """Dense P = P_tau (x) I_Vs from a cyclic 13-offset table."""
    Lt = table.shape[0]
    Pt = np.zeros((Lt, Lt))
    for tau in range(Lt):
        for o in range(-6, 7):
            Pt[tau, (tau + o) % Lt] += table[tau, o + 6]
    return np.kron(Pt, np.eye(Vs))


def equal_time_G(L, Lt, phi, dtau=0.1, t=1.0):
    """Diagonal blocks of M^{-1}: G_tau = (I + B_{tau-1} ... B_0 B_{Lt-1} ... B_tau)^{-1}, which
    follows from solving M x = e on the ring of slices (each block step multiplies by B, the
    antiperiodic corner adds one sign).  [Lt, Vs, Vs] complex."""
    Bs = np.stack([B_mat(L, phi[tau], dtau, t) for tau in range(Lt)])
    A = np.broadcast_to(np.eye(L * L, dtype=np.complex128), Bs.shape).copy()
    for k in range(Lt):
        A = Bs[(np.arange(Lt) + k) % Lt] @ A
    return np.linalg.inv(np.eye(L * L)[None] + A)


def bond_x(L, Lt, phi, dtau=0.1, t=1.0):
    """Exact noise average of the reference's x-bond field Tv (stochastic_estimator.py:1631-1654),
    from M^{-1}: with E[eta*_i (M^{-1} eta)_j] = M^{-1}[j, i],
      B(tau, y, x) = M^{-1}[r+x, r] e^{i phi_x(r)} + M^{-1}[r, r+x] e^{-i phi_x(r)},  r = (tau, y, x),
    averaged over tau: [Ly, Lx] complex (the reference's B_r is its real part).  Both entries lie
    in the diagonal block of slice tau (equal_time_G)."""
    G = equal_time_G(L, Lt, phi, dtau, t)
    yy, xx = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    r, rx = yy * L + xx, yy * L + (xx + 1) % L
    u = np.exp(1j * np.asarray(phi)[..., 0])
    return (G[:, rx, r] * u + G[:, r, rx] * np.conj(u)).mean(axis=0)
