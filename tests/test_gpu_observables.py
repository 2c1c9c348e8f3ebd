"""Bond-bond estimators on the GPU (qf_obs_* kernels + rocFFT) against the numpy oracle.

Tolerances: complex128 inputs 1e-10 relative (FP64 throughout, summation order
differs); complex64 inputs 2e-5 relative (FP32 fields and transforms, FP64
k-space sums).
"""
import math

import numpy as np
import pytest
import torch

from oracle import observables as OB

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _inputs(L, Lt, nrv, seed, dtype=torch.complex128):
    rng = np.random.default_rng(seed)
    N = L * L * Lt
    eta = ((rng.integers(0, 2, (nrv, N)) * 2 - 1) + 1j * (rng.integers(0, 2, (nrv, N)) * 2 - 1)) / math.sqrt(2)
    G = rng.standard_normal((nrv, N)) + 1j * rng.standard_normal((nrv, N))
    boson = rng.uniform(-np.pi, np.pi, (1, 2, L, L, Lt))
    return eta, G, boson


def _dev(eta, G, boson, dev, dtype):
    rd = torch.float32 if dtype == torch.complex64 else torch.float64
    return (torch.from_numpy(eta).to(dev, dtype), torch.from_numpy(G).to(dev, dtype), torch.from_numpy(boson).to(dev, rd))


@pytest.mark.parametrize("method", ["fused", "rocfft"])
@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-10), (torch.complex64, 2e-5)])
@pytest.mark.parametrize("L,Lt,nrv", [(4, 10, 5), (6, 12, 7)])
def test_estimators_match_oracle(dev, dtype, tol, L, Lt, nrv, method):
    from qed_fermion_amd.observables import BondBondEstimator
    print(f"[obs] L={L} Lt={Lt} nrv={nrv} {dtype} {method}", flush=True)
    eta, G, boson = _inputs(L, Lt, nrv, seed=L * 100 + nrv)
    if dtype == torch.complex64:   # the oracle sees the rounded inputs
        eta = eta.astype(np.complex64).astype(np.complex128)
        G = G.astype(np.complex64).astype(np.complex128)
        boson = boson.astype(np.float32).astype(np.float64)
    e_d, G_d, b_d = _dev(eta, G, boson, dev, dtype)
    est = BondBondEstimator(L, Lt, nrv, dtype=dtype, device=dev, method=method)
    assert est.method == method
    for name in ("T1", "T21", "T2", "T4"):
        got = est.T(name, e_d, G_d, b_d).cpu().numpy()
        want = OB.T(name, eta, G, boson, L, Lt)
        assert _rel(got, want) < tol, name
    tvv, tv = est.Tvv(e_d, G_d, b_d)
    wvv, wv = OB.Tvv(eta, G, boson, L, Lt)
    assert _rel(tvv.cpu().numpy(), wvv) < tol
    assert _rel(tv.cpu().numpy(), wv) < tol
    got = est.bond_bond(e_d, G_d, b_d)
    want = OB.bond_bond(eta, G, boson, L, Lt)
    for k in ("BB0_r", "BB_r", "B_r", "BB_k"):
        assert _rel(got[k].cpu().numpy(), want[k]) < tol, k


def test_chunking_does_not_change_the_result(dev):
    from qed_fermion_amd.observables import BondBondEstimator
    L, Lt, nrv = 6, 16, 9
    eta, G, boson = _inputs(L, Lt, nrv, seed=5)
    e_d, G_d, b_d = _dev(eta, G, boson, dev, torch.complex128)
    whole = BondBondEstimator(L, Lt, nrv, dtype=torch.complex128, device=dev, method="rocfft")
    small = BondBondEstimator(L, Lt, nrv, dtype=torch.complex128, device=dev, pair_chunk=4, nwg=7, method="rocfft")
    for name in ("T1", "T21", "T2", "T4"):
        a = whole.T(name, e_d, G_d, b_d).cpu().numpy()
        b = small.T(name, e_d, G_d, b_d).cpu().numpy()
        assert _rel(b, a) < 1e-12, name
    a, b = whole.Tvv(e_d, G_d, b_d), small.Tvv(e_d, G_d, b_d)
    assert _rel(b[0].cpu().numpy(), a[0].cpu().numpy()) < 1e-12


@pytest.mark.parametrize("L", [6, 10, 14, 22, 26, 36, 48])
def test_fused_transform_radices(dev, L):
    """The fused kernel's own FFT for lattices whose radices are 4, 2, 3, 5, 7 and larger primes
    (11, 13 take the generic stage), complex64 against the oracle."""
    from qed_fermion_amd.observables import BondBondEstimator
    Lt, nrv = 4, 4
    print(f"[obs] fused radices L={L}", flush=True)
    eta, G, boson = _inputs(L, Lt, nrv, seed=L)
    eta = eta.astype(np.complex64).astype(np.complex128)
    G = G.astype(np.complex64).astype(np.complex128)
    boson = boson.astype(np.float32).astype(np.float64)
    e_d, G_d, b_d = _dev(eta, G, boson, dev, torch.complex64)
    est = BondBondEstimator(L, Lt, nrv, dtype=torch.complex64, device=dev, method="fused")
    assert _rel(est.T("T1", e_d, G_d, b_d).cpu().numpy(), OB.T("T1", eta, G, boson, L, Lt)) < 2e-5
    tvv, tv = est.Tvv(e_d, G_d, b_d)
    wvv, wv = OB.Tvv(eta, G, boson, L, Lt)
    assert _rel(tvv.cpu().numpy(), wvv) < 2e-5 and _rel(tv.cpu().numpy(), wv) < 2e-5


def test_fused_unsupported_sizes_fall_back(dev):
    from qed_fermion_amd import _lib
    from qed_fermion_amd.observables import BondBondEstimator
    assert _lib.lib().qf_obs_fused_grid(_lib.QF_PREC_F64, 64) == 0      # complex128 slices of 64 x 64 exceed LDS
    assert _lib.lib().qf_obs_fused_grid(_lib.QF_PREC_F32, 64) > 0
    est = BondBondEstimator(64, 2, 2, dtype=torch.complex128, device=dev)
    assert est.method == "rocfft"


def test_fused_skips_malformed_pairs(dev):
    from qed_fermion_amd import _lib
    from qed_fermion_amd._lib import call, ptr, stream_handle
    L, Lt, nrv = 6, 8, 3
    N, V = L * L * Lt, L * L
    eta, G, boson = _inputs(L, Lt, nrv, seed=4)
    e_d, G_d, _ = _dev(eta, G, boson, dev, torch.complex128)
    U = torch.exp(1j * torch.from_numpy(boson[0, 0].transpose(2, 1, 0).copy())).to(dev)
    nwg = _lib.lib().qf_obs_fused_grid(_lib.QF_PREC_F64, L)
    part = torch.empty(2 * nwg, V, dtype=torch.complex128, device=dev)
    out = torch.empty(2, V, dtype=torch.complex128, device=dev)
    for i, pl in enumerate(([0, 1], [0, 1, 1, nrv, -1, 2])):
        pairs = torch.tensor(pl, dtype=torch.int32, device=dev)
        call("qf_obs_fused", _lib.QF_PREC_F64, _lib.QF_OBS_T1, ptr(e_d), ptr(G_d), ptr(U), ptr(pairs), len(pl) // 2, nrv, L, Lt,
             ptr(part), nwg, stream_handle())
        call("qf_obs_reduce", ptr(part), nwg, V, ptr(out[i]), stream_handle())
    torch.cuda.synchronize()
    assert _rel(out[1].cpu().numpy(), out[0].cpu().numpy()) < 1e-13


def test_malformed_pair_list_gives_zero_fields(dev):
    from qed_fermion_amd import _lib
    from qed_fermion_amd._lib import call, ptr, stream_handle
    L, Lt, nrv = 4, 6, 3
    N = L * L * Lt
    eta, G, boson = _inputs(L, Lt, nrv, seed=2)
    e_d, G_d, _ = _dev(eta, G, boson, dev, torch.complex128)
    U = torch.ones(N, dtype=torch.complex128, device=dev)
    pairs = torch.tensor([0, 1, 1, nrv, -1, 2], dtype=torch.int32, device=dev)   # pair 1 and 2 out of range
    A = torch.full((3, N), 7.0, dtype=torch.complex128, device=dev)
    B = torch.full_like(A, 7.0)
    call("qf_obs_fields", _lib.QF_PREC_F64, _lib.QF_OBS_T1, ptr(e_d), ptr(G_d), ptr(U), ptr(pairs), 0, 3, nrv, L, Lt,
         ptr(A), ptr(B), stream_handle())
    torch.cuda.synchronize()
    assert torch.count_nonzero(A[0]) > 0
    assert torch.count_nonzero(A[1:]) == 0 and torch.count_nonzero(B[1:]) == 0


def _sampler(dev, L, Lt, precision):
    from qed_fermion_amd.hmc_sampler import HmcSampler
    torch.manual_seed(0)
    return HmcSampler(Lx=L, Ltau=Lt, bs=2, precision=precision, device=dev)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fermion_obsr_pipeline_against_oracle(dev, precision):
    """get_fermion_obsr_compile: per chain G_eta solve, then the estimators; checked against
    the oracle evaluated on the same eta / G_eta / boson."""
    from qed_fermion_amd.stochastic_estimator import StochaticEstimator
    L, Lt, nrv = 6, 40, 6
    hmc = _sampler(dev, L, Lt, precision)
    se = StochaticEstimator(hmc, cuda_graph_se=False, Nrv=nrv, max_iter_se=300)
    se.initialize()
    eta = se.random_vec_bin()
    out = se.get_fermion_obsr_compile(hmc.boson, eta)
    assert out["BB_r"].shape == (2, L, L)
    tol = 1e-10 if precision == "fp64" else 2e-5
    # the last chain's G_eta is still held by the estimator
    e_np = eta.cpu().numpy().astype(np.complex128)
    G_np = se.G_eta.cpu().numpy().astype(np.complex128)
    b_np = hmc.boson[1:2].cpu().numpy().astype(np.float64)
    want = OB.bond_bond(e_np, G_np, b_np, L, Lt)
    for k in ("BB0_r", "BB_r", "B_r", "BB_k"):
        assert _rel(out[k][1].cpu().numpy(), want[k]) < tol, k


def test_estimator_graph_runners_replay(dev):
    """Captured T1..Tvv graphs replayed several times, with eager work between replays,
    each replay equal to the eager evaluation."""
    from qed_fermion_amd.stochastic_estimator import StochaticEstimator
    L, Lt, nrv = 6, 40, 6
    hmc = _sampler(dev, L, Lt, "fp64")
    se = StochaticEstimator(hmc, cuda_graph_se=True, Nrv=nrv, max_iter_se=200)
    se.initialize()
    se.init_cuda_graph()
    for rep in range(3):
        eta = se.random_vec_bin()
        boson = hmc.boson[rep % 2:rep % 2 + 1].clone()
        G = torch.randn(nrv, L * L * Lt, dtype=torch.complex128, device=dev)
        eager = {n: getattr(se, n)(eta, G, boson) for n in ("T1", "T21", "T2", "T4")}
        eager["Tvv"] = se.Tvv(eta, G, boson)
        torch.cuda.synchronize()
        for n, runner in (("T1", se.T1_graph_runner), ("T21", se.T21_graph_runner), ("T2", se.T2_graph_runner),
                          ("T4", se.T4_graph_runner)):
            got = runner(eta, G, boson).clone()
            torch.cuda.synchronize()
            assert _rel(got.cpu().numpy(), eager[n].cpu().numpy()) < 1e-13, (rep, n)
        tvv, tv = se.Tvv_graph_runner(eta, G, boson)
        assert _rel(tvv.cpu().numpy(), eager["Tvv"][0].cpu().numpy()) < 1e-13
        assert _rel(tv.cpu().numpy(), eager["Tvv"][1].cpu().numpy()) < 1e-13
