"""The drop-in Python API (reference-shaped classes and the _C shim) on the GPU,
against the oracle and the dense brute force."""
import math

import numpy as np
import pytest
import torch

import dense_ref as D
from helpers import make_phi, make_vec, relerr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sampler(L=4, Lt=30, bs=2, precision="fp64", **kw):
    from qed_fermion_amd.hmc_sampler import HmcSampler
    kw.setdefault("cuda_graph", False)
    return HmcSampler(Lx=L, Ltau=Lt, J=1.0, bs=bs, precision=precision, device="cuda", **kw)


def _ref_layout(phi_np):
    """oracle layout [bs, Lt, L, L, 2] -> reference boson [bs, 2, Lx, Ly, Ltau] (torch, cuda)."""
    return torch.from_numpy(phi_np).permute(0, 4, 3, 2, 1).contiguous().cuda()


def test_C_shim_matches_oracle(dev):
    from qed_fermion_amd import _C, precon
    L, Lt, bs = 6, 30, 2
    phi = make_phi(L, Lt, "uniform", seed=1, nchain=bs)
    x = make_vec(L, Lt, seed=2, nrhs=bs)
    boson = torch.from_numpy(phi.reshape(bs, -1)).float().cuda()
    xc = torch.from_numpy(x).to(torch.complex64).cuda()
    y = _C.mhm_vec(boson, xc, L, 0.1, 4, 8).cpu().numpy()
    z = _C.mh_vec(boson, xc, L, 0.1, 4, 8).cpu().numpy()
    for i in range(bs):
        assert relerr(y[i], O.apply_MdagM(phi[i], x[i], L, Lt)) < 2e-5
        assert relerr(z[i], O.apply_Mdag(phi[i], x[i], L, Lt)) < 2e-5
    csr = precon.coo_to_csr(precon.get_precon_man(L, Lt), device="cuda")
    w = _C.precon_vec(xc, csr, L).cpu().numpy()
    tab = O.precon_table(Lt)
    for i in range(bs):
        assert relerr(w[i], O.apply_P(tab, x[i], L * L, Lt)) < 1e-6
    # b_vec_per_tau on one slice: full B, and the intermediates in force_f_fast's order
    t = 3
    bt = boson[0].view(Lt, -1)[t]
    v = xc[0].view(Lt, -1)[t]
    full = _C.b_vec_per_tau(bt, v, L, 0.1, False).cpu().numpy()
    assert relerr(full, O.apply_B(phi[0, t], x[0].reshape(Lt, -1)[t], L)) < 2e-5
    inter = _C.b_vec_per_tau(bt, v, L, 0.1, True).cpu().numpy()
    for k, m in enumerate((5, 4, 3, 2, 1, 6)):
        assert relerr(inter[k], O.apply_B(phi[0, t], x[0].reshape(Lt, -1)[t], L, nfac=m)) < 2e-5


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-7), ("fp32", 2e-4)])
def test_force_f_fast_matches_oracle(dev, precision, tol):
    L, Lt, bs = 4, 30, 2
    hmc = _sampler(L, Lt, bs, precision, max_iter=800)
    hmc.cg_rtol = 1e-12 if precision != "fp32" else 1e-6
    phi = make_phi(L, Lt, "gauss", seed=5, nchain=bs)
    psi = make_vec(L, Lt, seed=6, nrhs=bs)
    boson = _ref_layout(phi).to(hmc.dtype)
    Ft, xi, iters, err = hmc.force_f_fast(torch.from_numpy(psi).to(hmc.cdtype).cuda(), boson, None)
    assert Ft.shape == (bs, 2, L, L, Lt) and xi.shape == (bs, L * L * Lt)
    tab = O.precon_table(Lt)
    for i in range(bs):
        xr, _, _ = O.pcg(phi[i], psi[i], L, Lt, table=tab, rtol=1e-13, max_iter=4000)
        Fr = O.force(phi[i], xr, L, Lt)                                      # [Lt, L, L, 2]
        Fr_ref_layout = np.transpose(Fr, (3, 2, 1, 0))                      # [2, Lx, Ly, Lt]
        assert relerr(Ft[i].cpu().numpy(), Fr_ref_layout) < tol
        assert relerr(xi[i].cpu().numpy(), xr) < tol


def test_boson_api_matches_oracle(dev):
    L, Lt, bs = 4, 10, 2
    hmc = _sampler(L, Lt, bs, "fp64")
    cJ, cK = hmc.couplings()
    phi = make_phi(L, Lt, "gauss", seed=8, nchain=bs)
    boson = _ref_layout(phi)
    for compact in (False, True):
        hmc.compact = compact
        S_tau = hmc.action_boson_tau(boson).cpu().numpy()
        S_pl = hmc.action_boson_plaq_matfree(boson).cpu().numpy()
        Fp = hmc.force_b_plaq_matfree(boson).cpu().numpy()
        Ftau = (hmc.force_b_tau_cmp if compact else hmc.force_b_tau_noncmp)(boson).cpu().numpy()
        for i in range(bs):
            assert abs(S_tau[i] - O.action_tau(phi[i], L, Lt, cJ, compact)) < 1e-10 * abs(S_tau[i]) + 1e-12
            assert abs(S_pl[i] - O.action_plaq(phi[i], L, Lt, cK)) < 1e-10 * abs(S_pl[i]) + 1e-12
            assert relerr(Fp[i], np.transpose(O.force_b_plaq(phi[i], L, Lt, cK), (3, 2, 1, 0))) < 1e-12
            assert relerr(Ftau[i], np.transpose(O.force_b_tau(phi[i], L, Lt, cJ, compact), (3, 2, 1, 0))) < 1e-12
    curl = hmc.curl_phi(boson)
    assert curl.shape == (bs, L, L, Lt)


def test_force_graph_runner_replays_force_f_fast(dev):
    from qed_fermion_amd.force_graph_runner import ForceGraphRunner
    L, Lt, bs = 6, 30, 2
    hmc = _sampler(L, Lt, bs, "fp32", max_iter=100)
    phi = make_phi(L, Lt, "gauss", seed=9, nchain=bs)
    boson = _ref_layout(phi).float()
    psi = torch.from_numpy(make_vec(L, Lt, seed=10, nrhs=bs)).to(torch.complex64).cuda()
    Ft0, xi0, _, err0 = hmc.force_f_fast(psi, boson, None)
    runner = ForceGraphRunner(hmc)
    runner.capture(torch.zeros_like(psi), torch.zeros_like(boson), 100)
    for _ in range(2):
        Ft, xi, err = runner(psi, boson)
        torch.cuda.synchronize()
        assert torch.equal(Ft, Ft0) and torch.equal(xi, xi0) and torch.equal(err, err0)


def test_leapfrog_cmp_and_runner(dev):
    from qed_fermion_amd.metropolis_graph_runner import LeapfrogCmpGraphRunner
    L, Lt, bs = 4, 10, 2
    hmc = _sampler(L, Lt, bs, "fp64")
    cJ, cK = hmc.couplings()
    phi = make_phi(L, Lt, "gauss", seed=11, nchain=bs)
    p = np.random.default_rng(12).standard_normal(phi.shape)
    x, pt = _ref_layout(phi), _ref_layout(p)
    dt = torch.tensor([0.05, 0.02], dtype=torch.float64, device="cuda").view(-1, 1, 1, 1, 1)
    mask = torch.ones(1, 1, 1, 1, Lt, dtype=torch.float64, device="cuda")
    Fp, Ft = hmc.force_b_plaq_matfree(x), hmc.force_b_tau_noncmp(x)
    x1, p1, Fp1, Ft1 = hmc.leapfrog_cmp(x, pt, dt, mask, Fp, Ft)
    for i in range(bs):
        xs, ps = phi[i].copy(), p[i].copy()
        h = [0.05, 0.02][i]
        fp, ft = O.force_b_plaq(xs, L, Lt, cK), O.force_b_tau(xs, L, Lt, cJ)
        for _ in range(5):
            ps = ps + (fp + ft) * h / 10
            xs = xs + ps * h / 5
            fp, ft = O.force_b_plaq(xs, L, Lt, cK), O.force_b_tau(xs, L, Lt, cJ)
            ps = ps + (fp + ft) * h / 10
        assert relerr(x1[i].cpu().numpy(), np.transpose(xs, (3, 2, 1, 0))) < 1e-13
        assert relerr(p1[i].cpu().numpy(), np.transpose(ps, (3, 2, 1, 0))) < 1e-13
    hmc.cuda_graph = True
    runner = LeapfrogCmpGraphRunner(hmc)
    runner.capture(torch.zeros_like(x), torch.zeros_like(x), torch.zeros_like(dt), mask.clone(), torch.zeros_like(x),
                   torch.zeros_like(x))
    out = runner(x, pt, dt, mask, Fp, Ft)
    torch.cuda.synchronize()
    for a, b in zip(out, (x1, p1, Fp1, Ft1)):
        assert torch.equal(a, b)


@pytest.mark.parametrize("cuda_graph", [False, True])
def test_metropolis_update_runs(dev, cuda_graph):
    L, Lt, bs = 6, 30, 3
    hmc = _sampler(L, Lt, bs, "mixed", max_iter=300, cuda_graph=cuda_graph)
    hmc.delta_t_tensor.fill_(0.02)
    for _ in range(2):
        before = hmc.boson.clone()
        boson, accp, iters, r_err = hmc.metropolis_update()
        torch.cuda.synchronize()
        assert boson.shape == (bs, 2, L, L, Lt)
        assert torch.equal(boson[~accp], before[~accp])
        assert torch.isfinite(r_err).all()
    assert all(len(q) == 2 for q in hmc.threshold_queue)
    assert all(0.0 <= v <= 1.0 for q in hmc.threshold_queue for v in q)


@pytest.mark.parametrize("cuda_graph", [False, True])
def test_tau_block_updates(dev, cuda_graph):
    """Two tau blocks (hmc_sampler_batch.py:2901-2925): a proposal with block 0's mask leaves
    block 1 of the field untouched, and metropolis_update runs one accept step per block."""
    L, Lt, bs = 4, 80, 2                      # Ltau >= 10 Lx, so blocks are allowed
    hmc = _sampler(L, Lt, bs, "fp64", max_iter=200, cuda_graph=cuda_graph, max_tau_block_idx=2)
    assert hmc.max_tau_block_idx == 2 and hmc.tau_block_size == 40
    assert hmc.threshold_queue[0].maxlen == 10
    hmc.delta_t_tensor.fill_(0.02)
    mask = torch.zeros((1, 1, 1, 1, Lt), device=dev, dtype=hmc.dtype)
    mask[..., :40] = 1.0
    x, H0, H1, _, _ = hmc.leapfrog_proposer5(hmc.boson, mask)
    torch.cuda.synchronize()
    assert torch.equal(x[..., 40:], hmc.boson[..., 40:])
    assert not torch.equal(x[..., :40], hmc.boson[..., :40])
    assert torch.isfinite(H0).all() and torch.isfinite(H1).all()
    boson, accp, iters, r_err = hmc.metropolis_update()
    torch.cuda.synchronize()
    assert all(len(q) == 2 for q in hmc.threshold_queue) and hmc.tau_block_idx == 0
    small = _sampler(4, 30, bs, "fp64", max_tau_block_idx=2)   # Ltau < 10 Lx: one block
    assert small.max_tau_block_idx == 1 and small.tau_block_size == 30


def test_z2_noise_is_orthonormal_on_average(dev):
    """random_vec_bin draws on the device (stochastic_estimator.py:210-224): every entry is one
    of (+-1 +- i)/sqrt(2), so |eta_i|^2 = 1 exactly, and the noise average of eta eta^dag is the
    identity up to O(1/sqrt(Nrv)) off the diagonal (the reference's orthogonality check,
    unit_test/test_stochastic_estimator.py:52-54), with the four phases equally likely."""
    from qed_fermion_amd.stochastic_estimator import StochaticEstimator
    L, Lt, nrv = 2, 30, 8192
    torch.manual_seed(5)
    hmc = _sampler(L, Lt, 1, "fp64")
    se = StochaticEstimator(hmc, cuda_graph_se=False, Nrv=nrv)
    e = se.random_vec_bin().to(torch.complex128).cpu().numpy()     # [Nrv, N]
    N = e.shape[1]
    assert np.allclose(np.abs(e) ** 2, 1.0, atol=1e-12)
    C = e.T @ e.conj() / nrv                                         # (1/Nrv) sum_s eta_s eta_s^dag
    assert np.allclose(np.diag(C), 1.0, atol=1e-12)
    off = np.abs(C - np.diag(np.diag(C)))
    assert off.max() < 6.0 / math.sqrt(nrv), off.max()
    assert abs(np.abs(C).sum() - N) / (N * (N - 1)) < 2.0 / math.sqrt(nrv)   # no systematic correlation
    ph = np.round(np.angle(e) / (np.pi / 4)).astype(int) % 8         # 1, 3, 5, 7: the four phases
    counts = np.bincount(ph.ravel(), minlength=8)
    assert counts[[0, 2, 4, 6]].sum() == 0
    frac = counts[[1, 3, 5, 7]] / ph.size
    assert np.all(np.abs(frac - 0.25) < 6 * math.sqrt(0.25 * 0.75 / ph.size)), frac


def test_stochastic_estimator_solve(dev):
    from qed_fermion_amd.stochastic_estimator import StochaticEstimator
    L, Lt = 4, 30
    hmc = _sampler(L, Lt, 1, "fp64", max_iter=1000)
    hmc.cg_rtol = 1e-12
    se = StochaticEstimator(hmc, cuda_graph_se=False, Nrv=8)
    se.initialize()
    eta = se.random_vec_bin()
    e = eta.cpu().numpy()
    assert np.allclose(np.abs(e.real), 1 / math.sqrt(2)) and np.allclose(np.abs(e.imag), 1 / math.sqrt(2))
    phi = make_phi(L, Lt, "gauss", seed=13)
    boson = _ref_layout(phi)
    se.set_eta_G_eta(boson, eta)
    M = D.M_mat(L, Lt, phi[0])
    G_ref = np.linalg.solve(M, e.T).T                      # G eta = M^{-1} eta
    assert relerr(se.G_eta.cpu().numpy(), G_ref) < 1e-9
    se.cuda_graph_se = True
    se.init_cuda_graph()
    assert se.G_eta_graph_runner.max_iter_se == hmc._MAX_ITERS_TO_CAPTURE[0] == 1000
    se.set_eta_G_eta(boson, eta)
    assert relerr(se.G_eta.cpu().numpy(), G_ref) < 1e-9


def test_geta_graph_runner_reference_call(dev):
    """GEtaGraphRunner called exactly as stochastic_estimator.py:135-141 does: capture(max_iter_se=...,
    graph_memory_pool=...) allocates its own static inputs, captures the solve with the sampler's
    max_iter set to max_iter_se and restores it; replays equal eager solves with that budget."""
    from qed_fermion_amd.fermion_obsr_graph_runner import GEtaGraphRunner
    from qed_fermion_amd.stochastic_estimator import StochaticEstimator
    L, Lt = 4, 30
    hmc = _sampler(L, Lt, 1, "fp64", max_iter=400)
    hmc.cg_rtol = 1e-12
    se = StochaticEstimator(hmc, cuda_graph_se=True, Nrv=6)
    se.initialize()
    runner = GEtaGraphRunner(se)
    pool = runner.capture(max_iter_se=12, graph_memory_pool=None)
    assert pool is not None and hmc.max_iter == 400
    assert runner.input_buffers["boson"].shape == (1, 2, L, L, Lt) and runner.input_buffers["eta"].shape == (6, L * L * Lt)
    for rep in range(3):
        phi = make_phi(L, Lt, "gauss", seed=30 + rep)
        boson = _ref_layout(phi)
        eta = se.random_vec_bin()
        got = runner(boson, eta).clone()
        torch.cuda.synchronize()
        hmc.max_iter = 12                                  # the captured budget, eagerly
        want = se.set_eta_G_eta_inner(boson, eta)
        hmc.max_iter = 400
        torch.cuda.synchronize()
        assert torch.equal(got, want), rep
        # 12 iterations leave this field unconverged: the budget was really the captured one
        full = se.set_eta_G_eta_inner(boson, eta)
        assert relerr(got.cpu().numpy(), full.cpu().numpy()) > 1e-8


def test_z2_noise_orthogonality(dev):
    """E[eta eta^dag] = I for the (+-1 +- i)/sqrt(2) noise (stochastic_estimator.py:264-289)."""
    from qed_fermion_amd.stochastic_estimator import StochaticEstimator
    hmc = _sampler(4, 30, 1, "fp64")
    se = StochaticEstimator(hmc, Nrv=4000)
    eta = se.random_vec_bin()[:, :16].to(torch.complex128)
    C = (eta.T @ eta.conj()) / eta.shape[0]
    assert torch.allclose(torch.diagonal(C).real, torch.ones(16, dtype=torch.float64, device="cuda"), atol=1e-12)
    off = C - torch.diag(torch.diagonal(C))
    assert off.abs().max().item() < 0.1


@pytest.mark.parametrize("cg_dtype", [torch.complex128, torch.complex64])
def test_cg_benchmark_matches_oracle(dev, cg_dtype):
    """CgHmcSampler.benchmark called as cg_hmc_sampler.py:256-341 does: the reference's 9-tuple,
    fields 6 and 4 only (in that order), and per field the HIP PCG's iteration counter and last
    residual against oracle.pcg at rtol 1e-7, max_iter 200, with and without the preconditioner."""
    from qed_fermion_amd.cg_hmc_sampler import CgHmcSampler
    L, Lt, nb = 4, 40, 5
    cg = CgHmcSampler(J=0.5, Nstep=3000, config=None, Lx=L, Ltau=Lt, bs=1, precision="fp32")
    cg.cg_dtype = cg_dtype
    cg.reset()
    cg.reset_precon()
    out = cg.benchmark(bs=nb)
    assert len(out) == 9
    assert out[0] is None and out[1] == -1.0 and out[2] is None and out[3] == []
    assert out[4] is None and out[5] is None and out[6] == [] and out[7] == []
    rec = cg.last_benchmark
    assert rec["fields"] == [6, 4] and out[8] == rec["err_plain"] and len(out[8]) == 2
    fields = cg.benchmark_fields(nb).cpu().numpy()
    assert fields.shape == (2 * nb, 2, L, L, Lt)
    b = rec["b"].to(torch.complex128).cpu().numpy()
    tab = O.precon_table(Lt)
    for k, i in enumerate(rec["fields"]):
        phi = np.ascontiguousarray(np.transpose(fields[i], (3, 2, 1, 0)))   # [Lt, Ly, Lx, 2]
        if cg_dtype == torch.complex64:
            phi = phi.astype(np.float32).astype(np.float64)                  # the field the FP32 kernels see
        for key_it, key_err, table in (("iters_precon", "err_precon", tab), ("iters_plain", "err_plain", None)):
            _, it_r, err_r = O.pcg(phi, b, L, Lt, table=table, rtol=1e-7, max_iter=200)
            it, err = rec[key_it][k], rec[key_err][k]
            # counters: the same recurrence agrees up to a threshold crossing in FP64, within a few
            # iterations in complex64; residuals: converged ones below rtol, unconverged ones (the
            # 200-iteration budget) equal up to the rounding the iteration amplifies
            slack = 1 if cg_dtype == torch.complex128 else max(3, int(0.05 * it_r))
            assert abs(it - it_r) <= slack, (i, key_it, it, it_r)
            if it_r < 200 and it < 200:
                assert err <= 1e-7 and err_r <= 1e-7, (i, key_err, err, err_r)
            else:   # complex64 iterates drift from FP64 ones by about 10% in 200 iterations on a rough field
                band = 0.05 if cg_dtype == torch.complex128 else 0.25
                assert abs(err - err_r) <= band * err_r, (i, key_err, err, err_r)
    # field 6 is pi flux plus the narrowest noise, the field family the preconditioner was fitted to
    assert rec["iters_precon"][0] < rec["iters_plain"][0]


def test_autograd_force(dev):
    """FermionAction.backward is -Ft; checked against the oracle force and central differences."""
    from qed_fermion_amd.autograd import fermion_action
    from qed_fermion_amd import kernels
    L, Lt = 4, 30
    phi = make_phi(L, Lt, "gauss", seed=14)
    psi = make_vec(L, Lt, seed=15)
    ptab = kernels.precon_table(Lt, device="cuda")
    ph = torch.from_numpy(phi.reshape(1, -1)).cuda().requires_grad_(True)
    ps = torch.from_numpy(psi).cuda()
    S = fermion_action(ph, ps, L, Lt, ptab=ptab, rtol=1e-14, max_iter=3000)
    S.sum().backward()
    xr, _, _ = O.pcg(phi[0], psi[0], L, Lt, table=O.precon_table(Lt), rtol=1e-14, max_iter=4000)
    assert relerr(ph.grad[0].cpu().numpy(), -O.force(phi[0], xr, L, Lt).ravel()) < 1e-9
    eps = 1e-5
    for k in np.random.default_rng(16).choice(phi.size, 6, replace=False):
        d = torch.zeros_like(ph)
        d.view(-1)[k] = eps
        with torch.no_grad():
            Sp = fermion_action(ph + d, ps, L, Lt, ptab=ptab, rtol=1e-14, max_iter=3000).item()
            Sm = fermion_action(ph - d, ps, L, Lt, ptab=ptab, rtol=1e-14, max_iter=3000).item()
        fd = (Sp - Sm) / (2 * eps)
        g = ph.grad.view(-1)[k].item()
        assert abs(fd - g) < 1e-5 * max(1.0, abs(g)), (k, fd, g)
    ph2 = ph.detach().clone().requires_grad_(True)
    fermion_action(ph2, ps, L, Lt, ptab=ptab, rtol=1e-14, max_iter=3000, recompute=False).sum().backward()
    assert torch.equal(ph2.grad, ph.grad)


def test_adaptive_mass_methods_on_device(dev):
    """The sampler's adaptive-mass methods (hmc_sampler_batch.py:1133-1277) on the device equal the
    same utilities on CPU tensors."""
    from qed_fermion_amd import fourier_accel as FA
    from qed_fermion_amd.hmc_sampler import HmcSampler
    torch.manual_seed(3)
    hmc = HmcSampler(Lx=4, Ltau=40, bs=2, precision="fp64", device=dev)
    ref = FA.SigmaHatState(2, 4, 4, 40, torch.device("cpu"), torch.float64, sigma_mini_batch_size=3)
    hmc._fa.sigma_mini_batch_size = 3
    for i in range(4):
        b = hmc.boson.cpu() + 0.1 * i * torch.randn(hmc.boson.shape, dtype=torch.float64)
        hmc.update_sigma_hat_cpu(b, i)
        FA.update_sigma_hat(ref, b, i)
    assert torch.allclose(hmc.sigma_hat_cpu, ref.sigma_hat_cpu, rtol=1e-12, atol=1e-14)
    hmc.apply_sigma_hat_cpu(3)
    FA.apply_sigma_hat(ref, 3)
    assert hmc.sigma_hat.is_cuda and torch.allclose(hmc.sigma_hat.cpu(), ref.sigma_hat, rtol=1e-10)
    p = hmc.draw_momentum_fft()
    assert p.shape == (2, 2, 4, 4, 40) and p.is_cuda
    ref.sigma_hat = hmc.sigma_hat.cpu()
    assert torch.allclose(hmc.apply_m_inv(p).cpu(), FA.apply_m_inv(ref, p.cpu()), rtol=1e-10, atol=1e-12)


def test_wrappers_reject_wrong_buffers(dev):
    """Output and workspace buffers of the wrong dtype, size or layout raise QfError before any
    kernel runs (instead of an out-of-bounds write or a write into a temporary copy)."""
    from qed_fermion_amd import kernels as K
    L, Lt = 4, 30
    N = L * L * Lt
    phi = torch.zeros(1, 2 * N, dtype=torch.float64, device=dev)
    x = torch.zeros(1, N, dtype=torch.complex128, device=dev)
    bad_outs = [torch.zeros(1, N - 1, dtype=torch.complex128, device=dev),          # short
                torch.zeros(1, N, dtype=torch.complex64, device=dev),               # dtype
                torch.zeros(2, N, dtype=torch.complex128, device=dev)[:, ::2]]      # strided view
    for out in bad_outs:
        with pytest.raises(K.QfError):
            K.apply("MdagM", phi, x, L, Lt, out=out)
        with pytest.raises(K.QfError):
            K.fermion_force(phi, x, L, Lt, out=torch.zeros(1, 2 * N - 2, dtype=torch.float64, device=dev))
    with pytest.raises(K.QfError):
        K.pcg_solve(phi, x, L, Lt, err=torch.zeros(1, dtype=torch.float32, device=dev))
    with pytest.raises(K.QfError):
        K.pcg_solve(phi, x, L, Lt, iters=torch.zeros(2, dtype=torch.float64, device=dev))
    with pytest.raises(K.QfError):
        K.pcg_solve(phi, x, L, Lt, work=torch.zeros(16, dtype=torch.uint8, device=dev))
    with pytest.raises(K.QfError):
        K.pcg_solve(phi.float(), x, L, Lt)                                          # phi / b dtypes differ
    with pytest.raises(K.QfError):
        K.boson_force(phi, L, Lt, 1.0, 1.0, Fp=torch.zeros(2 * N, dtype=torch.float32, device=dev))
    with pytest.raises(K.QfError):
        K.hamiltonian(phi, L, Lt, 1.0, 1.0, H=torch.zeros(3, dtype=torch.float64, device=dev))


def test_force_gathered_over_the_c_shim(dev):
    """The reference's own force assembly over the _C shim (Level-1 drop-in): force_f_fast
    (hmc_sampler_batch.py:2170-2227) gathers df_dot_bs1 (1987-2006) products of
    _C.b_vec_per_tau partial products (interm=True) per slice.  Restated here over this build's
    shim and checkerboard lists, it must give the fused adjoint force kernel's result for the
    same xi: this pins the shim's order and conjugation convention of the partial products,
    which no reference output shows.  The assembly pairs partial k with partial 4 - k, so it cannot
    tell the order 0..4 from its reverse; two partials swapped fail it."""
    from qed_fermion_amd import _C
    from qed_fermion_amd import kernels as K
    L, Lt, dtau, t = 4, 10, 0.1, 1.0
    hmc = _sampler(L, Lt, 1, "fp64")
    Vs = L * L
    phi = torch.from_numpy(make_phi(L, Lt, "gauss", seed=95).reshape(1, -1)).to(dev)
    xi = torch.from_numpy(make_vec(L, Lt, seed=96)).to(dev)
    bos = phi.view(Lt, 2 * Vs)
    xt = xi.view(Lt, Vs)

    def df_dot(blist, lft, rgt, il, jl, group1=False):
        v = (math.sinh(t * dtau) if group1 else math.sinh(t * dtau / 2)) * torch.exp(1j * (blist + math.pi / 2))
        return lft[il] * rgt[jl] * v + lft[jl] * rgt[il] * v.conj()

    def assemble(partials_order):
        F = torch.empty(Lt, 2 * Vs, dtype=torch.float64, device=dev)
        groups = [(hmc.boson_idx_list_1, hmc.i_list_1, hmc.j_list_1), (hmc.boson_idx_list_2, hmc.i_list_2, hmc.j_list_2),
                  (hmc.boson_idx_list_3, hmc.i_list_3, hmc.j_list_3), (hmc.boson_idx_list_4, hmc.i_list_4, hmc.j_list_4)]
        for tau in range(Lt):
            b = bos[tau]
            xc, xn = xt[tau], xt[(tau + 1) % Lt]
            part = lambda v: _C.b_vec_per_tau(b, v, L, dtau, True)[partials_order]
            Bx = _C.b_vec_per_tau(b, xc, L, dtau, False)
            nl = part(xn.conj())          # xi_n_lft_conj
            cr = part(xc)                 # xi_c_rgt
            bc = part(Bx.conj())          # B_xi_conj
            s = -1.0 if tau < Lt - 1 else 1.0
            (i1, il1, jl1), (i2, il2, jl2), (i3, il3, jl3), (i4, il4, jl4) = groups
            F[tau, i1] = 2 * torch.real(df_dot(b[i1], nl[2].conj(), cr[2], il1, jl1, True) * s
                                        + df_dot(b[i1], bc[2].conj(), cr[2], il1, jl1, True))
            F[tau, i2] = 2 * torch.real((df_dot(b[i2], nl[1].conj(), cr[3], il2, jl2) + df_dot(b[i2], nl[3].conj(), cr[1], il2, jl2)) * s
                                        + df_dot(b[i2], bc[1].conj(), cr[3], il2, jl2) + df_dot(b[i2], bc[3].conj(), cr[1], il2, jl2))
            F[tau, i3] = 2 * torch.real((df_dot(b[i3], nl[0].conj(), cr[4], il3, jl3) + df_dot(b[i3], nl[4].conj(), cr[0], il3, jl3)) * s
                                        + df_dot(b[i3], bc[0].conj(), cr[4], il3, jl3) + df_dot(b[i3], bc[4].conj(), cr[0], il3, jl3))
            F[tau, i4] = 2 * torch.real((df_dot(b[i4], xn.conj(), cr[5], il4, jl4) + df_dot(b[i4], nl[5].conj(), xc, il4, jl4)) * s
                                        + df_dot(b[i4], Bx.conj(), cr[5], il4, jl4) + df_dot(b[i4], bc[5].conj(), xc, il4, jl4))
        return F.view(1, -1)

    Fk = K.fermion_force(phi, xi, L, Lt, dtau, t)
    Fg = assemble(slice(None))
    scale = Fk.abs().max().item()
    assert scale > 0
    # the reference returns +2 Re(...), which is -dS_f/dphi up to the sign of its derivative convention
    err = min((Fg - Fk).abs().max().item(), (Fg + Fk).abs().max().item()) / scale
    assert err < 1e-10, err
    sign = 1.0 if (Fg - Fk).abs().max().item() < (Fg + Fk).abs().max().item() else -1.0
    print(f"[shim force] max rel diff {err:.2e}, sign {sign:+.0f}", flush=True)
    same = assemble(torch.tensor([4, 3, 2, 1, 0, 5], device=dev))    # the reverse order: indistinguishable
    assert (same - sign * Fk).abs().max().item() / scale < 1e-10
    wrong = assemble(torch.tensor([1, 0, 2, 3, 4, 5], device=dev))   # two partials swapped
    assert (wrong - sign * Fk).abs().max().item() / scale > 1e-3
