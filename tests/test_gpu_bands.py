"""Lattices too wide for one work-group per slice (L = 68, 70, 74, 76, 80, 96, ...): each slice
is split into bands of rows, every work-group recomputing a halo of 4 rows (one B_tau) or 8
rows (M^dag M, the force) on both sides.  The operator is defined for any even L
(hmc_sampler_batch.py:937-969, 1741-1906); every band result is checked against the oracle at
the bars of the whole-slice kernels."""
import numpy as np
import pytest
import torch

from helpers import make_phi, make_vec, relerr
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPS = {"M": O.apply_M, "Mdag": O.apply_Mdag, "MdagM": O.apply_MdagM}


def _k():
    from qed_fermion_amd import kernels
    return kernels


def _t(a, dev, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype)


@pytest.mark.parametrize("L", [68, 70, 74, 76, 80, 96])
@pytest.mark.parametrize("op", ["M", "Mdag", "MdagM", "B"])
def test_band_operator_f64(dev, L, op):
    """FP64 M, M^dag, M^dag M and B_tau at 1e-12 of the oracle; per-chain fields, two right-hand
    sides, every tau-run length 1..5 (a run edge in every slice position)."""
    k = _k()
    Lt, nrhs = 6, 2
    print(f"[band operator f64] L={L} op={op}", flush=True)
    phi = make_phi(L, Lt, "uniform", seed=L + 3, nchain=nrhs)
    x = make_vec(L, Lt, seed=L + 4, nrhs=nrhs)
    if op == "B":
        ref = [np.concatenate([O.apply_B(phi[r][t], x[r].reshape(Lt, -1)[t], L) for t in range(Lt)]) for r in range(nrhs)]
    else:
        ref = [OPS[op](phi[r], x[r], L, Lt) for r in range(nrhs)]
    pd, xd = _t(phi.reshape(nrhs, -1), dev, torch.float64), _t(x, dev, torch.complex128)
    try:
        for T in (1, 2, 5):
            k.set_tuning(1, 0, T, 0)
            y = k.apply(op, pd, xd, L, Lt).cpu().numpy()
            for r in range(nrhs):
                assert relerr(y[r], ref[r]) < 1e-12, (T, r, relerr(y[r], ref[r]))
    finally:
        k.set_tuning(1, 0, 0, 0)


@pytest.mark.parametrize("L", [70, 96])
def test_band_operator_f32_and_hop(dev, L):
    """complex64 M^dag M from phi and from the hop table, and FP64 M^dag M from the hop table."""
    k = _k()
    Lt, nrhs = 8, 2
    print(f"[band operator f32/hop] L={L}", flush=True)
    phi = make_phi(L, Lt, "uniform", seed=L + 5, nchain=nrhs)
    x = make_vec(L, Lt, seed=L + 6, nrhs=nrhs)
    ref = [O.apply_MdagM(phi[r], x[r], L, Lt) for r in range(nrhs)]
    for rd, cd, tol in ((torch.float32, torch.complex64, 2e-5), (torch.float64, torch.complex128, 1e-12)):
        pd, xd = _t(phi.reshape(nrhs, -1), dev, rd), _t(x, dev, cd)
        outs = [k.apply_hop("MdagM", k.hop_table(pd, L, Lt), xd, L, Lt)]
        if rd == torch.float32:
            outs.append(k.apply("MdagM", pd, xd, L, Lt))
        for y in outs:
            yn = y.cpu().numpy()
            for r in range(nrhs):
                assert relerr(yn[r], ref[r]) < tol, (rd, r, relerr(yn[r], ref[r]))


@pytest.mark.parametrize("L", [70, 80])
def test_band_partial_B(dev, L):
    """The partial products Q_m of B_tau (_C.b_vec_per_tau) on bands, m = 1..7."""
    k = _k()
    Lt = 4
    phi = make_phi(L, Lt, "uniform", seed=L + 7)
    x = make_vec(L, Lt, seed=L + 8)
    pd, xd = _t(phi.reshape(1, -1), dev, torch.float64), _t(x, dev, torch.complex128)
    for m in range(1, 8):
        y = k.apply_partial_B(pd, xd, L, Lt, 0.1, m).cpu().numpy().reshape(-1)
        ref = np.concatenate([O.apply_B(phi[0][t], x[0].reshape(Lt, -1)[t], L, nfac=m) for t in range(Lt)])
        assert relerr(y, ref) < 1e-12, (m, relerr(y, ref))


@pytest.mark.parametrize("L", [68, 70, 76, 80, 96])
def test_band_force_f64(dev, L):
    """The adjoint fermion force on bands at 1e-12 of the oracle."""
    k = _k()
    Lt, nrhs = 4, 2
    print(f"[band force f64] L={L}", flush=True)
    phi = make_phi(L, Lt, "uniform", seed=L + 9, nchain=nrhs)
    xi = make_vec(L, Lt, seed=L + 10, nrhs=nrhs)
    F = k.fermion_force(_t(phi.reshape(nrhs, -1), dev, torch.float64), _t(xi, dev, torch.complex128), L, Lt).cpu().numpy()
    for i in range(nrhs):
        ref = O.force(phi[i], xi[i], L, Lt)
        assert relerr(F[i], ref) < 1e-12, (i, relerr(F[i], ref))


def test_band_pcg_f64_and_mixed(dev):
    """PCG at L=70 (two bands): FP64 follows the oracle's iterates at 1e-10 with the reference's
    iteration counter; the mixed solver reaches the FP64 true residual recomputed by the oracle
    and agrees with FP64 within 1e-7."""
    k = _k()
    L, Lt, nrhs = 70, 30, 2
    print("[band pcg] L=70", flush=True)
    phi = make_phi(L, Lt, "gauss", seed=71, nchain=nrhs)
    R = make_vec(L, Lt, seed=72, nrhs=nrhs)
    tab = O.precon_table(Lt)
    b = np.stack([O.apply_Mdag(phi[i], R[i], L, Lt) for i in range(nrhs)])
    pd = _t(phi.reshape(nrhs, -1), dev, torch.float64)
    bd = _t(b, dev, torch.complex128)
    ptab = _t(tab, dev, torch.float64)
    x, iters, err = k.pcg_solve(pd, bd, L, Lt, ptab=ptab, rtol=1e-10, max_iter=300)
    xm, itm, errm = k.pcg_mixed_solve(pd, bd, L, Lt, ptab=ptab, rtol=1e-10, max_iter=300)
    x, xm = x.cpu().numpy(), xm.cpu().numpy()
    for i in range(nrhs):
        xr, it, er = O.pcg(phi[i], b[i], L, Lt, table=tab, rtol=1e-10, max_iter=300, early_exit=True)
        assert int(iters[i].item()) == it and er <= 1e-10
        assert relerr(x[i], xr) < 1e-9, relerr(x[i], xr)
        res = relerr(O.apply_MdagM(phi[i], xm[i], L, Lt), b[i])
        assert res <= 1.5e-10 and errm[i].item() <= 1e-10, (res, errm[i].item())
        assert relerr(xm[i], x[i]) < 1e-7


def test_fp64_residual_1024_thread_work_group(dev):
    """The mixed solver's FP64 true residual in its 1024-thread form (2x2 patches at L=64, whose
    FP64 tiling otherwise picks 2x4): it once held b in registers across a whole tau step and,
    spilled, came back wrong in the last step of every run.  The solve must reach the FP64
    residual recomputed by the oracle."""
    k = _k()
    L, Lt, nrhs = 64, 30, 1
    print("[fp64 residual 1024 threads] L=64", flush=True)
    phi = make_phi(L, Lt, "gauss", seed=81, nchain=nrhs)
    R = make_vec(L, Lt, seed=82, nrhs=nrhs)
    tab = O.precon_table(Lt)
    b = np.stack([O.apply_Mdag(phi[i], R[i], L, Lt) for i in range(nrhs)])
    k.set_tuning(1, 2, 0, 0)
    try:
        assert k.get_tiling(1, nrhs, L, Lt)["by"] == 2
        xm, itm, errm = k.pcg_mixed_solve(_t(phi.reshape(nrhs, -1), dev, torch.float64), _t(b, dev, torch.complex128), L, Lt,
                                          ptab=_t(tab, dev, torch.float64), rtol=1e-10, max_iter=300)
        xm = xm.cpu().numpy()
    finally:
        k.set_tuning(1, 0, 0, 0)
    res = relerr(O.apply_MdagM(phi[0], xm[0], L, Lt), b[0])
    assert res <= 1.5e-10 and errm[0].item() <= 1e-10 and itm[0].item() < 300, (res, errm[0].item())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_band_trajectory(dev, precision):
    """A whole HMC trajectory on a banded lattice (L=70, Ltau=30, 3 leapfrog steps, every solve
    to 1e-10) against the oracle's trajectory: final field and momenta within 1e-8 (FP64) or 1e-7
    (mixed), H0 to 1e-9 and Delta H to 1e-6."""
    from qed_fermion_amd.trajectory import TrajectoryEngine
    L, Lt, nl = 70, 30, 3
    print(f"[band trajectory] {precision}", flush=True)
    phi = make_phi(L, Lt, "gauss", seed=7001)
    p = np.random.default_rng(7002).standard_normal(phi.shape)
    R = make_vec(L, Lt, seed=7003)
    eng = TrajectoryEngine(L, Lt, 1, precision=precision, max_iter=1500, rtol=1e-10, n_leap=nl, delta_t=0.05, device="cuda")
    eng.set_phi(torch.from_numpy(phi.reshape(1, -1)).cuda())
    eng.p.copy_(torch.from_numpy(p.reshape(1, -1)))
    eng.R.copy_(torch.from_numpy(R))
    eng.propose(draw=False)
    torch.cuda.synchronize()
    x1, p1, H0, H1, err = O.trajectory(phi[0], p[0], R[0], L, Lt, dt=0.05, n_leap=nl, table=O.precon_table(Lt), rtol=1e-10,
                                       max_iter=1500)
    tol = 1e-8 if precision == "fp64" else 1e-7
    assert eng.err.max().item() <= 1e-10 and err <= 1e-10
    assert relerr(eng.x[0].cpu().numpy(), x1) < tol
    assert relerr(eng.p[0].cpu().numpy(), p1) < tol
    assert abs(eng.H0[0].item() - H0) < 1e-9 * abs(H0)
    assert abs((eng.H1[0] - eng.H0[0]).item() - (H1 - H0)) < 1e-6
