"""Seeded synthetic lattices shared by the tests (numpy, oracle layout)."""
import numpy as np


def make_phi(L, Lt, kind="uniform", seed=0, nchain=1):
    """phi [nchain, Lt, L, L, 2] float64: 'uniform' in [-pi, pi) or 'gauss' with sigma 0.3."""
    rng = np.random.default_rng(seed)
    shape = (nchain, Lt, L, L, 2)
    if kind == "uniform":
        return rng.uniform(-np.pi, np.pi, shape)
    if kind == "gauss":
        return 0.3 * rng.standard_normal(shape)
    if kind == "piflux":
        phi = np.zeros(shape)
        for x in range(L):
            for y in range(L):
                s = (-1) ** (x + y)
                phi[:, :, y, x, 0] = s * np.pi / 4
                phi[:, :, y, x, 1] = -s * np.pi / 4
        return phi
    raise ValueError(kind)


def make_vec(L, Lt, seed=1, nrhs=1):
    rng = np.random.default_rng(seed)
    N = L * L * Lt
    return (rng.standard_normal((nrhs, N)) + 1j * rng.standard_normal((nrhs, N))) / np.sqrt(2)


def relerr(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
