"""The plain-method configurations end to end against the oracle:
C1 -- L=6, Ltau=60, phi uniform, complex128 PCG to 1e-10, 20 leapfrog steps and
Metropolis; C2 -- L=24, Ltau=240, phi ~ N(0, 0.3^2), complex128 PCG to 1e-10, the
40-step trajectory captured as one hipGraph."""
import numpy as np
import pytest
import torch

from helpers import make_phi, make_vec, relerr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _load(eng, phi, p, R):
    eng.set_phi(torch.from_numpy(phi.reshape(phi.shape[0], -1)).cuda())
    eng.p.copy_(torch.from_numpy(p.reshape(p.shape[0], -1)))
    eng.R.copy_(torch.from_numpy(R))


def test_c1_trajectory_and_metropolis(dev):
    from qed_fermion_amd.trajectory import TrajectoryEngine
    L, Lt, nl = 6, 60, 20
    print("[C1]", flush=True)
    phi = make_phi(L, Lt, "uniform", seed=601)
    p = np.random.default_rng(602).standard_normal(phi.shape)
    R = make_vec(L, Lt, seed=603)
    eng = TrajectoryEngine(L, Lt, 1, precision="fp64", max_iter=3000, rtol=1e-10, n_leap=nl, delta_t=0.06, device="cuda")
    _load(eng, phi, p, R)
    eng.propose(draw=False)
    torch.cuda.synchronize()
    x1, p1, H0, H1, err = O.trajectory(phi[0], p[0], R[0], L, Lt, dt=0.06, n_leap=nl, table=O.precon_table(Lt), rtol=1e-10,
                                       max_iter=3000)
    assert eng.err.max().item() <= 1e-10 and err <= 1e-10
    assert relerr(eng.x[0].cpu().numpy(), x1) < 1e-8
    assert relerr(eng.p[0].cpu().numpy(), p1) < 1e-8
    assert abs(eng.H0[0].item() - H0) < 1e-9 * abs(H0)
    assert abs((eng.H1[0] - eng.H0[0]).item() - (H1 - H0)) < 1e-6
    # Metropolis: accept iff u < exp(H0 - H1); a rejected chain returns to phi exactly
    x_prop = eng.x.clone()
    eng.metropolis()
    acc = eng.u[0].item() < np.exp(eng.H0[0].item() - eng.H1[0].item())
    assert bool(eng.accept[0].item()) == acc
    want = x_prop if acc else torch.from_numpy(phi.reshape(1, -1)).cuda()
    assert torch.equal(eng.x, want)


def test_c2_graph_trajectory(dev):
    from qed_fermion_amd.trajectory import TrajectoryEngine
    L, Lt, nl = 24, 240, 40
    print("[C2]", flush=True)
    phi = make_phi(L, Lt, "gauss", seed=2401)
    p = np.random.default_rng(2402).standard_normal(phi.shape)
    R = make_vec(L, Lt, seed=2403)
    eng = TrajectoryEngine(L, Lt, 1, precision="fp64", max_iter=1500, rtol=1e-10, n_leap=nl, delta_t=0.02, device="cuda")
    _load(eng, phi, p, R)
    eng.propose(draw=False)
    eager = [t.clone() for t in (eng.x, eng.p, eng.H0, eng.H1)]
    _load(eng, phi, p, R)
    eng.capture(n_warmup=1, draw=False, accept=False)
    for rep in range(2):
        _load(eng, phi, p, R)
        torch.cuda.synchronize()
        eng.replay()
        torch.cuda.synchronize()
        for a, b in zip((eng.x, eng.p, eng.H0, eng.H1), eager):
            assert torch.equal(a, b), rep
    assert eng.err.max().item() <= 1e-10
    x1, p1, H0, H1, err = O.trajectory(phi[0], p[0], R[0], L, Lt, dt=0.02, n_leap=nl, table=O.precon_table(Lt), rtol=1e-10,
                                       max_iter=1500)
    assert relerr(eng.x[0].cpu().numpy(), x1) < 1e-8
    assert relerr(eng.p[0].cpu().numpy(), p1) < 1e-8
    assert abs((eng.H1[0] - eng.H0[0]).item() - (H1 - H0)) < 1e-6


def test_c2_graph_trajectory_at_1e12(dev):
    """C2 at the configuration's own bar: the 40-step FP64 trajectory replayed from its hipGraph,
    every solve converged to 1e-14, ends within 1e-12 relative of the oracle's in field and
    momenta (every 64th entry and 16 Gaussian projections, tests/golden/make_c2_trajectory.py),
    with H0 and H1 within 1e-12 relative.  At 1e-14 the solver's rounding moves the oracle's own
    final state by less than the bar."""
    import os
    from qed_fermion_amd.trajectory import TrajectoryEngine
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_trajectory_rtol1e-14.npz"))
    L, Lt, nl, stride, nproj, pseed = (int(v) for v in d["meta"])
    print("[C2 1e-12]", flush=True)
    phi = make_phi(L, Lt, "gauss", seed=2401)
    p = np.random.default_rng(2402).standard_normal(phi.shape)
    R = make_vec(L, Lt, seed=2403)
    eng = TrajectoryEngine(L, Lt, 1, precision="fp64", max_iter=4000, rtol=float(d["rtol"]), n_leap=nl,
                           delta_t=float(d["dt"]), device="cuda")
    _load(eng, phi, p, R)
    eng.propose(draw=False)                 # eager: the iterations each solve needs
    torch.cuda.synchronize()
    need = eng.iters[:, 0].tolist()
    print(f"[C2 1e-12] iterations per solve {min(need):.0f}-{max(need):.0f}", flush=True)
    assert (eng.status == 0).all()
    eng.max_iter = [int(v) + 20 for v in need]   # caps per solve, as the trajectory's gears
    _load(eng, phi, p, R)
    eng.capture(n_warmup=1, draw=False, accept=False)
    _load(eng, phi, p, R)
    eng.replay()
    torch.cuda.synchronize()
    assert (eng.status == 0).all() and eng.err.max().item() <= 1e-14
    x = eng.x[0].cpu().numpy().ravel()
    pm = eng.p[0].cpu().numpy().ravel()
    assert relerr(x[::stride], d["x_sub"]) < 1e-12
    assert relerr(pm[::stride], d["p_sub"]) < 1e-12
    G = np.random.default_rng(pseed).standard_normal((nproj, x.size))
    assert np.abs(G @ x - d["x_proj"]).max() < 1e-12 * float(d["x_norm"])
    assert np.abs(G @ pm - d["p_proj"]).max() < 1e-12 * float(d["p_norm"])
    H0, H1 = float(d["H0"]), float(d["H1"])
    assert abs(eng.H0[0].item() - H0) < 1e-12 * abs(H0)
    assert abs(eng.H1[0].item() - H1) < 1e-12 * abs(H1)
