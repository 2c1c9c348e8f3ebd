"""Per-chain solve status of the HIP PCG solvers (QF_SOLVE_* codes): a NaN in phi or b, a
breakdown, a fixed budget.  Failed chains stop at their last finite iterate and the other
chains of the batch are unaffected (checked against the oracle)."""
import numpy as np
import pytest
import torch

from helpers import make_phi, make_vec, relerr
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CONV, MAXIT, NONFIN, BRK = 0, 1, 2, 3


def _solve(prec, phi, b, L, Lt, ptab, rtol, max_iter):
    from qed_fermion_amd import kernels as K
    nrhs = b.shape[0]
    status = torch.full((nrhs,), -1, dtype=torch.int32, device="cuda")
    if prec == "mixed":
        x, it, err = K.pcg_mixed_solve(phi, b, L, Lt, ptab=ptab, rtol=rtol, max_iter=max_iter, ru_period=50, status=status)
    else:
        r = torch.float64 if prec == "fp64" else torch.float32
        c = torch.complex128 if prec == "fp64" else torch.complex64
        x, it, err = K.pcg_solve(phi.to(r), b.to(c), L, Lt, ptab=ptab, rtol=rtol, max_iter=max_iter, status=status)
    torch.cuda.synchronize()
    return x, it, err, status


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_nan_phi_and_nan_b_are_flagged(dev, prec):
    from qed_fermion_amd import kernels as K
    L, Lt, n = 4, 30, 4
    phi = make_phi(L, Lt, "gauss", seed=40, nchain=n)
    b = make_vec(L, Lt, seed=41, nrhs=n)
    phi[1, 3, 1, 2, 0] = np.nan     # one link of chain 1
    b[2, 17] = np.nan               # one entry of chain 2's right-hand side
    ptab = K.precon_table(Lt, device=dev)
    rtol = 1e-10 if prec != "fp32" else 1e-6
    x, it, err, status = _solve(prec, torch.from_numpy(phi.reshape(n, -1)).to(dev), torch.from_numpy(b).to(dev), L, Lt,
                                ptab, rtol, 600)
    assert status.tolist() == [CONV, NONFIN, NONFIN, CONV]
    assert torch.isfinite(x).all(), "a failed chain's NaN reached x"
    # err is the last iterated residual; the mixed solver reports the FP64 true residual, which is
    # NaN for a NaN field
    assert torch.isfinite(err[[0, 3] if prec == "mixed" else [0, 1, 3]]).all()
    tab = O.precon_table(Lt)
    tol = 1e-9 if prec != "fp32" else 1e-4
    for i in (0, 3):   # the healthy chains of the batch
        xr, _, _ = O.pcg(phi[i], b[i], L, Lt, table=tab, rtol=1e-13, max_iter=4000)
        assert relerr(x[i].cpu().numpy(), xr) < tol, i


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_breakdown_and_max_iter(dev, prec):
    """A zero preconditioner makes z = p = 0, so Re(p^dag A p) = 0 at the first step: a breakdown
    with x left at 0.  A budget too small for rtol ends with QF_SOLVE_MAX_ITER."""
    L, Lt, n = 4, 30, 2
    phi = torch.from_numpy(make_phi(L, Lt, "gauss", seed=42, nchain=n).reshape(n, -1)).to(dev)
    b = torch.from_numpy(make_vec(L, Lt, seed=43, nrhs=n)).to(dev)
    zero = torch.zeros(Lt, 13, dtype=torch.float64, device=dev)
    x, it, err, status = _solve(prec, phi, b, L, Lt, zero, 1e-10, 50)
    assert status.tolist() == [BRK, BRK]
    assert torch.count_nonzero(x).item() == 0 and torch.isfinite(err).all()
    from qed_fermion_amd import kernels as K
    x, it, err, status = _solve(prec, phi, b, L, Lt, K.precon_table(Lt, device=dev), 1e-10, 3)
    assert status.tolist() == [MAXIT, MAXIT] and (err > 1e-10).all()
    assert it.tolist() == [3.0, 3.0]


def test_trajectory_rejects_failed_chains(dev):
    """Inside the captured trajectory the Metropolis step rejects chains whose solves failed:
    with a zero preconditioner every solve breaks down (x = 0, finite forces and H), and
    every chain returns to its start although H stays finite."""
    from qed_fermion_amd.trajectory import TrajectoryEngine
    L, Lt, bs = 4, 30, 3
    eng = TrajectoryEngine(L, Lt, bs, precision="fp64", max_iter=50, rtol=1e-10, n_leap=2, delta_t=0.02, device=dev)
    phi0 = torch.from_numpy(make_phi(L, Lt, "gauss", seed=44, nchain=bs).reshape(bs, -1)).to(dev)
    eng.set_phi(phi0)
    eng.ptab.zero_()
    eng.capture(n_warmup=1)
    for _ in range(2):
        eng.set_phi(phi0)
        eng.replay()
        torch.cuda.synchronize()
        assert (eng.status == BRK).all()
        assert torch.isfinite(eng.H0).all() and torch.isfinite(eng.H1).all()
        assert not eng.accept.any()
        assert torch.equal(eng.x, phi0)


def test_sampler_warns_and_rejects(dev):
    """HmcSampler.metropolis_update rejects a chain with a NaN field (its solves are non-finite)
    and warns; the healthy chain proceeds."""
    from qed_fermion_amd.hmc_sampler import HmcSampler
    hmc = HmcSampler(Lx=4, Ltau=30, J=1.0, bs=2, precision="fp64", device="cuda", cuda_graph=False, max_iter=300)
    hmc.delta_t_tensor.fill_(0.02)
    hmc.boson[0, 0, 1, 1, 5] = float("nan")
    with pytest.warns(RuntimeWarning, match="rejected"):
        boson, accp, _, _ = hmc.metropolis_update()
    torch.cuda.synchronize()
    assert not accp[0].item()
    assert (hmc.last_solve_status[:, 0] == NONFIN).all() and (hmc.last_solve_status[:, 1] == CONV).all()
