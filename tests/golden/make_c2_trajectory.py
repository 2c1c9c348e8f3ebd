"""Writes tests/golden/c2_trajectory_rtol1e-14.npz: the oracle's 40-step FP64 trajectory of the
L=24, Ltau=240 configuration (phi ~ N(0, 0.3^2) per bond, seeds 2401-2403 of tests/helpers.py),
every solve converged to 1e-14, reduced to what a 1e-12 comparison needs: every 64th entry of
the final field and momenta, 16 Gaussian projections of each (seeded, regenerated by the test),
their norms and the energies.  Solving to 1e-14 moves the final state by 3e-13 (x) and 1e-12 (p)
relative against 1e-13 solves, so the state is fixed to the 1e-12 bar by the field and noise,
not by the solver's rounding.  About three minutes on 8 host threads:

    python tests/golden/make_c2_trajectory.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from helpers import make_phi, make_vec  # noqa: E402
from oracle import oracle as O  # noqa: E402

L, LT, NLEAP, DT, RTOL, MAXIT = 24, 240, 40, 0.02, 1e-14, 4000
STRIDE, NPROJ, PROJ_SEED = 64, 16, 2404


def projections(n):
    return np.random.default_rng(PROJ_SEED).standard_normal((NPROJ, n))


def main():
    phi = make_phi(L, LT, "gauss", seed=2401)
    p = np.random.default_rng(2402).standard_normal(phi.shape)
    R = make_vec(L, LT, seed=2403)
    x1, p1, H0, H1, err = O.trajectory(phi[0], p[0], R[0], L, LT, dt=DT, n_leap=NLEAP, table=O.precon_table(LT),
                                       rtol=RTOL, max_iter=MAXIT)
    x1, p1 = x1.ravel(), p1.ravel()
    G = projections(x1.size)
    np.savez(os.path.join(HERE, "c2_trajectory_rtol1e-14.npz"),
             x_sub=x1[::STRIDE], p_sub=p1[::STRIDE], x_proj=G @ x1, p_proj=G @ p1,
             x_norm=np.linalg.norm(x1), p_norm=np.linalg.norm(p1), H0=H0, H1=H1, err=err,
             meta=np.array([L, LT, NLEAP, STRIDE, NPROJ, PROJ_SEED], dtype=np.int64), dt=DT, rtol=RTOL)
    print(f"H0 {H0:.12e}  dH {H1 - H0:.6e}  last residual {err:.3e}")


if __name__ == "__main__":
    main()
