"""Statistical agreement of the GPU sampler with the oracle's: the captured trajectory
(leapfrog_proposer5 + metropolis_update, hmc_sampler_batch.py:2366-2625, 2886-2930, with its
on-device draws of p and R and its on-device Metropolis step) run under the protocol of
tests/golden/make_hmc_statistics.py, whose oracle chains give the expected means.  The
deterministic parity tests fix the trajectory for given (phi, p, R); this checks what they
leave open: the distributions of the draws and the accept/reject step, through the run
averages of S_tau, S_plaq (of the current field, evaluated by the oracle on both sides), the
acceptance and a fermion observable, B_x: the mean x-bond of the reference's Tv/B_r
(stochastic_estimator.py:1631-1654) in its exact noise average from M^{-1} of the current field
(tests/dense_ref.py bond_x, evaluated the same way on both sides every fermion_every-th measured
trajectory), each within 4 combined standard errors over independent chains."""
import json
import math
import os

import numpy as np
import pytest
import torch

import dense_ref as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sampler_statistics_match_the_oracle(dev, precision):
    """fp64: the FP64 sampler; mixed: the headline's complex64 PCG with FP64 reliable updates
    (every solve still to the FP64 true residual rtol), against the same FP64 oracle chains.
    With the same seed the mixed chains draw the same p, R and u and, their Delta H within
    about 1e-8 of FP64, take the same decisions: their statistics coincide with the FP64 run's
    (measured to every printed digit), so this also shows the mixed sampler tracking the FP64
    one over 64 x 400 trajectories."""
    from qed_fermion_amd.trajectory import TrajectoryEngine
    ref = json.load(open(os.path.join(HERE, "golden", "hmc_statistics_L4_Lt30.json")))
    P = ref["protocol"]
    L, Lt, bs = P["L"], P["Ltau"], 64
    torch.manual_seed(31)
    print(f"[statistics] {precision}", flush=True)
    eng = TrajectoryEngine(L, Lt, bs, precision=precision, dtau=P["dtau"], J=P["J"], K=P["K"], compact=P["compact"],
                           m=P["m"], delta_t=P["dt"], n_leap=P["n_leap"], max_iter=150, rtol=P["rtol"], device=dev)
    eng.set_phi(torch.zeros(bs, 2 * L * L * Lt, dtype=torch.float64, device=dev))
    eng.capture(n_warmup=1)
    eng.set_phi(torch.zeros(bs, 2 * L * L * Lt, dtype=torch.float64, device=dev))
    cJ, cK = O.couplings(P["J"], P["K"], P["dtau"])
    sums = np.zeros((bs, 4))
    nf = 0
    worst = 0.0
    for k in range(P["n_therm"] + P["n_meas"]):
        eng.replay()
        if k >= P["n_therm"]:
            x = eng.x.cpu().numpy().reshape(bs, Lt, L, L, 2)
            acc = eng.accept.cpu().numpy()
            worst = max(worst, eng.err.max().item())
            ferm = (k - P["n_therm"]) % P["fermion_every"] == 0
            nf += ferm
            for c in range(bs):
                sums[c, :3] += (O.action_tau(x[c], L, Lt, cJ, P["compact"]), O.action_plaq(x[c], L, Lt, cK), float(acc[c]))
                if ferm:
                    sums[c, 3] += D.bond_x(L, Lt, x[c], dtau=P["dtau"]).real.mean()
    torch.cuda.synchronize()
    assert (eng.status == 0).all() and worst <= P["rtol"], worst    # every solve converged within its 150 iterations
    means = np.concatenate([sums[:, :3] / P["n_meas"], sums[:, 3:] / nf], axis=1)
    for i, name in enumerate(("S_tau", "S_plaq", "acceptance", "B_x")):
        m = means[:, i].mean()
        se = means[:, i].std(ddof=1) / math.sqrt(bs)
        r = ref[name]
        z = (m - r["mean"]) / math.sqrt(se ** 2 + r["se"] ** 2)
        print(f"[statistics] {name}: GPU {m:.4f} +- {se:.4f}, oracle {r['mean']:.4f} +- {r['se']:.4f}, z = {z:+.2f}", flush=True)
        assert abs(z) < 4.0, (name, m, r["mean"], z)
