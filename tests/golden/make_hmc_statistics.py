"""Writes tests/golden/hmc_statistics_L4_Lt30.json: equilibrium-run statistics of the oracle's
HMC (leapfrog_proposer5 + metropolis_update restated in oracle/qf_oracle.c, with the
reference's draws: p ~ N(0, m), R = randn(complex)/sqrt(2)) on L=4, Ltau=30, for comparison
with the GPU sampler run under the same protocol (tests/test_gpu_statistics.py).  Per chain:
the means over the measured trajectories of S_tau and S_plaq of the current field and of the
acceptance, and over every fermion_every-th measured trajectory of the fermion observable B_x:
the mean over sites of the reference's x-bond field B_r (stochastic_estimator.py:1631-1654) in
its exact noise average from M^{-1} (tests/dense_ref.py bond_x).  64 chains from phi = 0, 8 worker processes with one OpenMP thread each; about two
minutes on 8 host cores:

    python tests/golden/make_hmc_statistics.py
"""
import json
import math
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

PROTOCOL = {"L": 4, "Ltau": 30, "dtau": 0.1, "J": 1.0, "K": 1.0, "compact": False, "m": 1.0, "dt": 0.1, "n_leap": 5,
            "rtol": 1e-10, "max_iter": 400, "start": "phi = 0", "n_therm": 100, "n_meas": 300, "fermion_every": 5}
CHAINS, SEED = 64, 2024


def run_chain(c):
    os.environ["OMP_NUM_THREADS"] = "1"
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.dirname(HERE))
    import numpy as np
    import dense_ref as D
    from oracle import oracle as O
    P = PROTOCOL
    L, Lt = P["L"], P["Ltau"]
    N = L * L * Lt
    cJ, cK = O.couplings(P["J"], P["K"], P["dtau"])
    tab = O.precon_table(Lt)
    rng = np.random.default_rng([SEED, c])
    phi = np.zeros((Lt, L, L, 2))
    sums = np.zeros(4)
    nf = 0
    for k in range(P["n_therm"] + P["n_meas"]):
        p = rng.standard_normal(phi.shape) * math.sqrt(P["m"])
        R = (rng.standard_normal(N) + 1j * rng.standard_normal(N)) / 2.0
        x1, _, H0, H1, _ = O.trajectory(phi, p, R, L, Lt, dtau=P["dtau"], J=P["J"], K=P["K"], compact=P["compact"],
                                        m=P["m"], dt=P["dt"], n_leap=P["n_leap"], table=tab, rtol=P["rtol"],
                                        max_iter=P["max_iter"])
        acc = rng.uniform() < math.exp(min(0.0, H0 - H1))
        if acc:
            phi = x1
        if k >= P["n_therm"]:
            sums[:3] += (O.action_tau(phi, L, Lt, cJ, P["compact"]), O.action_plaq(phi, L, Lt, cK), float(acc))
            if (k - P["n_therm"]) % P["fermion_every"] == 0:
                sums[3] += D.bond_x(L, Lt, phi, dtau=P["dtau"]).real.mean()
                nf += 1
    return (sums[:3] / P["n_meas"]).tolist() + [sums[3] / nf]


def main():
    with mp.get_context("spawn").Pool(8) as pool:
        chains = pool.map(run_chain, range(CHAINS))
    names = ("S_tau", "S_plaq", "acceptance", "B_x")
    out = {"protocol": PROTOCOL, "chains": CHAINS, "seed": SEED, "chain_means": {n: [c[i] for c in chains] for i, n in enumerate(names)}}
    for i, n in enumerate(names):
        v = [c[i] for c in chains]
        m = sum(v) / len(v)
        se = math.sqrt(sum((x - m) ** 2 for x in v) / (len(v) - 1) / len(v))
        out[n] = {"mean": m, "se": se}
        print(f"{n}: {m:.4f} +- {se:.4f}")
    with open(os.path.join(HERE, "hmc_statistics_L4_Lt30.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
