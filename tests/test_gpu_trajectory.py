"""GPU trajectory engine against the oracle's leapfrog_proposer5 restatement, and
hipGraph replays against eager execution."""
import math

import numpy as np
import pytest
import torch

from helpers import make_phi, make_vec, relerr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _engine(L, Lt, bs, precision, **kw):
    from qed_fermion_amd.trajectory import TrajectoryEngine
    return TrajectoryEngine(L, Lt, bs, precision=precision, device="cuda", **kw)


def _to_fermion_layout(phi):   # [bs, Lt, L, L, 2] numpy -> [bs, 2N]
    return torch.from_numpy(phi.reshape(phi.shape[0], -1)).cuda()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-7)])
def test_trajectory_matches_oracle(dev, precision, tol):
    L, Lt, bs = 4, 30, 2
    print(f"[trajectory] {precision}", flush=True)
    phi = make_phi(L, Lt, "gauss", seed=61, nchain=bs)
    p = np.random.default_rng(62).standard_normal(phi.shape)
    R = make_vec(L, Lt, seed=63, nrhs=bs)
    eng = _engine(L, Lt, bs, precision, max_iter=600, rtol=1e-12, n_leap=3, delta_t=0.03)
    eng.set_phi(_to_fermion_layout(phi))
    eng.p.copy_(torch.from_numpy(p.reshape(bs, -1)))
    eng.R.copy_(torch.from_numpy(R))
    eng.propose(draw=False)
    torch.cuda.synchronize()
    tab = O.precon_table(Lt)
    for i in range(bs):
        x1, p1, H0, H1, err = O.trajectory(phi[i], p[i], R[i], L, Lt, dt=0.03, n_leap=3, table=tab, rtol=1e-12,
                                           max_iter=600)
        assert relerr(eng.x[i].cpu().numpy(), x1) < tol
        assert relerr(eng.p[i].cpu().numpy(), p1) < tol
        assert abs(eng.H0[i].item() - H0) < tol * abs(H0)
        assert abs((eng.H1[i] - eng.H0[i]).item() - (H1 - H0)) < 1e-6 + tol * abs(H0)
    assert eng.err.max().item() < (1e-11 if precision == "fp64" else 2e-12)


def test_graph_replay_matches_eager(dev):
    """Replays of the captured trajectory (draw=False, accept=False) reproduce eager execution
    on every replay, with eager work and synchronisations in between."""
    L, Lt, bs = 6, 30, 2
    phi = make_phi(L, Lt, "gauss", seed=71, nchain=bs)
    p = np.random.default_rng(72).standard_normal(phi.shape)
    R = make_vec(L, Lt, seed=73, nrhs=bs)
    eng = _engine(L, Lt, bs, "mixed", max_iter=200, rtol=1e-10, n_leap=2, delta_t=0.03)

    def load():
        eng.set_phi(_to_fermion_layout(phi))
        eng.p.copy_(torch.from_numpy(p.reshape(bs, -1)))
        eng.R.copy_(torch.from_numpy(R))

    load()
    eng.propose(draw=False)
    ref = [t.clone() for t in (eng.x, eng.p, eng.H0, eng.H1, eng.xi)]
    load()
    eng.capture(n_warmup=1, draw=False, accept=False)
    for rep in range(3):
        load()
        torch.cuda.synchronize()
        eng.replay()
        torch.cuda.synchronize()
        got = (eng.x, eng.p, eng.H0, eng.H1, eng.xi)
        for a, b in zip(got, ref):
            assert torch.equal(a, b), rep
        _ = torch.randn(1000, device="cuda").sum().item()   # unrelated eager work between replays


def test_full_graph_trajectory_runs_and_accepts(dev):
    """Full captured trajectory with on-device draws and Metropolis: finite energies,
    acceptance in [0,1], rejected chains restored exactly."""
    L, Lt, bs = 6, 30, 4
    phi = make_phi(L, Lt, "gauss", seed=81, nchain=bs)
    eng = _engine(L, Lt, bs, "mixed", max_iter=300, rtol=1e-9, n_leap=2, delta_t=0.02)
    eng.set_phi(_to_fermion_layout(phi))
    eng.capture(n_warmup=1)
    for _ in range(3):
        before = eng.x.clone()
        eng.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(eng.H0).all() and torch.isfinite(eng.H1).all()
        acc = eng.accept
        assert torch.equal(eng.x[~acc], before[~acc])
        dH = (eng.H1 - eng.H0).abs().max().item()
        assert dH < 50.0
