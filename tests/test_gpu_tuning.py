"""Autotuner: tuned tilings are registered per shape and keep results exact."""
import numpy as np
import pytest
import torch

from helpers import make_phi, make_vec, relerr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_tune_operator_and_pcg(dev, tmp_path, monkeypatch):
    from qed_fermion_amd import autotune, kernels
    monkeypatch.setenv("QF_TUNING_CACHE", str(tmp_path / "tuning.json"))
    L, Lt, nrhs = 12, 40, 3
    print("[autotune] operator", flush=True)
    try:
        hit = autotune.tune(0, L, Lt, nrhs, kind="operator", reps=3)
        assert hit["by"] in autotune.valid_by(L, 0)
        til = kernels.get_tiling(0, nrhs, L, Lt)
        assert (til["by"], til["t"]) == (hit["by"], hit["t"])
        phi = make_phi(L, Lt, "uniform", seed=1)
        x = make_vec(L, Lt, seed=2, nrhs=nrhs)
        hop = kernels.hop_table(torch.from_numpy(phi.reshape(1, -1)).to("cuda", torch.float32), L, Lt)
        y = kernels.apply_hop("MdagM", hop, torch.from_numpy(x).to("cuda", torch.complex64), L, Lt).cpu().numpy()
        for i in range(nrhs):
            assert relerr(y[i], O.apply_MdagM(phi[0], x[i], L, Lt)) < 2e-5
        print("[autotune] operator from phi, per-chain fields", flush=True)
        hit3 = autotune.tune(0, L, Lt, nrhs, kind="operator", shared_field=False, source="phi", reps=3)
        assert hit3["ns"] in autotune.ns_candidates(L, hit3["by"], nrhs)
        assert kernels.get_tiling(0, nrhs, L, Lt)["ns"] == hit3["ns"]
        php = make_phi(L, Lt, "uniform", seed=5, nchain=nrhs)
        y = kernels.apply("MdagM", torch.from_numpy(php.reshape(nrhs, -1)).to("cuda", torch.float32),
                          torch.from_numpy(x).to("cuda", torch.complex64), L, Lt).cpu().numpy()
        for i in range(nrhs):
            assert relerr(y[i], O.apply_MdagM(php[i], x[i], L, Lt)) < 2e-5
        print("[autotune] pcg", flush=True)
        hit2 = autotune.tune(1, L, Lt, 2, kind="pcg", iters=5, reps=1)
        assert kernels.get_tiling(1, 2, L, Lt)["tb"] == hit2["tb"] or hit2["tb"] == 0
        # cached: a second call reuses the entry
        assert autotune.tune(1, L, Lt, 2, kind="pcg") == hit2
        kernels.clear_tuning()
        assert autotune.load() == 3
        phd = torch.from_numpy(make_phi(L, Lt, "gauss", seed=3, nchain=2).reshape(2, -1)).to("cuda", torch.float64)
        R = torch.from_numpy(make_vec(L, Lt, seed=4, nrhs=2)).to("cuda", torch.complex128)
        b = kernels.apply("Mdag", phd, R, L, Lt)
        xs, _, err = kernels.pcg_solve(phd, b, L, Lt, ptab=kernels.precon_table(Lt), rtol=1e-10, max_iter=500)
        r = b - kernels.apply("MdagM", phd, xs, L, Lt)
        assert (torch.linalg.vector_norm(r, dim=1) / torch.linalg.vector_norm(b, dim=1)).max().item() < 1.1e-10
    finally:
        kernels.clear_tuning()
