"""GPU parity of the HIP kernels against the FP64 CPU oracle.

Bars: FP64 kernels within 1e-12 relative (operator, force, boson sector);
FP32 kernels within 2e-5 relative; PCG in FP64 follows the oracle's iterates
within 1e-8 relative (same recurrence, different reduction order); the mixed
solver reaches the requested FP64 true residual and stays within 1e-7 of the
FP64 solution.
"""
import numpy as np
import pytest
import torch

from helpers import make_phi, make_vec, relerr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _k():
    from qed_fermion_amd import kernels
    return kernels


def _t(a, dev, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype)


OPS = {"M": O.apply_M, "Mdag": O.apply_Mdag, "MdagM": O.apply_MdagM}


@pytest.mark.parametrize("L,Lt,nrhs,shared", [(2, 4, 1, False), (4, 10, 3, False), (6, 12, 2, True), (12, 30, 2, False),
                                              (36, 8, 1, False)])
@pytest.mark.parametrize("op", ["M", "Mdag", "MdagM", "B"])
@pytest.mark.parametrize("by", [0, 4, 6])
def test_operator_f64(dev, L, Lt, nrhs, shared, op, by):
    k = _k()
    if by and L % by:
        pytest.skip("patch height does not divide L")
    print(f"[operator f64] L={L} Lt={Lt} nrhs={nrhs} shared={shared} op={op} by={by}", flush=True)
    k.set_tuning(1, by, 0, 0)
    try:
        phi = make_phi(L, Lt, "uniform", seed=L + Lt, nchain=1 if shared else nrhs)
        x = make_vec(L, Lt, seed=7, nrhs=nrhs)
        y = k.apply(op, _t(phi.reshape(phi.shape[0], -1), dev, torch.float64), _t(x, dev, torch.complex128), L, Lt)
        y = y.cpu().numpy()
        for r in range(nrhs):
            ph = phi[0 if shared else r]
            if op == "B":
                ref = np.concatenate([O.apply_B(ph[t], x[r].reshape(Lt, -1)[t], L) for t in range(Lt)])
            else:
                ref = OPS[op](ph, x[r], L, Lt)
            assert relerr(y[r], ref) < 1e-12, (r, relerr(y[r], ref))
    finally:
        k.set_tuning(1, 0, 0, 0)


@pytest.mark.parametrize("L,Lt,nrhs", [(4, 10, 2), (24, 24, 1)])
@pytest.mark.parametrize("op", ["Mdag", "MdagM"])
def test_operator_f32(dev, L, Lt, nrhs, op):
    k = _k()
    print(f"[operator f32] L={L} Lt={Lt} op={op}", flush=True)
    phi = make_phi(L, Lt, "uniform", seed=3, nchain=nrhs)
    x = make_vec(L, Lt, seed=4, nrhs=nrhs)
    y = k.apply(op, _t(phi.reshape(nrhs, -1), dev, torch.float32), _t(x, dev, torch.complex64), L, Lt).cpu().numpy()
    for r in range(nrhs):
        ref = OPS[op](phi[r], x[r], L, Lt)
        assert relerr(y[r], ref) < 2e-5


@pytest.mark.parametrize("L,Lt,nrhs", [(4, 30, 2), (6, 60, 1), (12, 40, 3)])
def test_precon(dev, L, Lt, nrhs):
    k = _k()
    print(f"[precon] L={L} Lt={Lt}", flush=True)
    tab = O.precon_table(Lt)
    r = make_vec(L, Lt, seed=5, nrhs=nrhs)
    z = k.apply_precon(_t(tab, dev, torch.float64), _t(r, dev, torch.complex128), L * L, Lt).cpu().numpy()
    for i in range(nrhs):
        assert relerr(z[i], O.apply_P(tab, r[i], L * L, Lt)) < 1e-13
    z32 = k.apply_precon(_t(tab, dev, torch.float32), _t(r, dev, torch.complex64), L * L, Lt).cpu().numpy()
    assert relerr(z32[0], O.apply_P(tab, r[0], L * L, Lt)) < 1e-6


@pytest.mark.parametrize("L,Lt,kind,precon,max_iter,xtol", [(4, 30, "gauss", True, 60, 1e-10),
                                                            (6, 60, "uniform", True, 150, 1e-5),
                                                            (6, 12, "uniform", False, 80, 1e-6)])
def test_pcg_f64_follows_oracle(dev, L, Lt, kind, precon, max_iter, xtol):
    """Same recurrence as the oracle in FP64.  CG iterates amplify reduction-order rounding
    (the oracle itself moves by ~1e-3 in the residual when only its summation order changes), so
    iterates are compared at xtol (1e-10 on a well-conditioned field, looser on the ill-conditioned
    uniform one), and the reported error must be the true residual of the returned x."""
    k = _k()
    print(f"[pcg f64] L={L} Lt={Lt} {kind} precon={precon}", flush=True)
    nrhs = 2
    phi = make_phi(L, Lt, kind, seed=11, nchain=nrhs)
    R = make_vec(L, Lt, seed=12, nrhs=nrhs)
    tab = O.precon_table(Lt) if precon else None
    b = np.stack([O.apply_Mdag(phi[i], R[i], L, Lt) for i in range(nrhs)])
    x, iters, err = k.pcg_solve(_t(phi.reshape(nrhs, -1), dev, torch.float64), _t(b, dev, torch.complex128), L, Lt,
                                ptab=_t(tab, dev, torch.float64) if precon else None, rtol=1e-9, max_iter=max_iter)
    x = x.cpu().numpy()
    for i in range(nrhs):
        xr, it, er = O.pcg(phi[i], b[i], L, Lt, table=tab, rtol=1e-9, max_iter=max_iter, early_exit=False)
        assert relerr(x[i], xr) < xtol, relerr(x[i], xr)
        true_res = relerr(O.apply_MdagM(phi[i], x[i], L, Lt), b[i])
        assert abs(err[i].item() - true_res) < 1e-2 * true_res + 1e-13, (err[i].item(), true_res)
        assert abs(err[i].item() - er) <= max(100 * xtol, 1e-3) * er + 1e-14
        assert int(iters[i].item()) == it


def test_pcg_converged_chain_freezes(dev):
    """A chain that reaches rtol stops updating (the active mask of 802-803, 841)."""
    k = _k()
    L, Lt = 4, 30
    phi = make_phi(L, Lt, "gauss", seed=2, nchain=2)
    b = make_vec(L, Lt, seed=3, nrhs=2)
    tab = O.precon_table(Lt)
    x, iters, err = k.pcg_solve(_t(phi.reshape(2, -1), dev, torch.float64), _t(b, dev, torch.complex128), L, Lt,
                                ptab=_t(tab, dev, torch.float64), rtol=1e-6, max_iter=300)
    for i in range(2):
        xr, it, er = O.pcg(phi[i], b[i], L, Lt, table=tab, rtol=1e-6, max_iter=300, early_exit=True)
        assert int(iters[i].item()) == it < 299
        assert relerr(x.cpu().numpy()[i], xr) < 1e-7
        # the residual at the stopping iteration is as sensitive to rounding as the
        # iterates themselves; both stop on the same iteration below rtol
        assert err[i].item() <= 1e-6 and er <= 1e-6


@pytest.mark.parametrize("prec,rtol,max_iter", [("f64", 1e-10, 600), ("f64", 0.0, 37), ("f32", 1e-5, 600),
                                              ("mixed", 1e-10, 600), ("mixed", 1e-10, 37)])
def test_paired_x_updates_equal_stepwise(dev, prec, rtol, max_iter):
    """x updates held and applied in pairs (the default) give the x of one update per
    iteration (qf_set_paired_x_updates(0)) bit for bit, with the same iteration counts:
    chains that stop right after a held iteration (even counts) or after a full one (odd
    counts), a solve that ends on a held iteration (max_iter odd), and the mixed solver's
    reliable updates, which read x every few iterations."""
    from qed_fermion_amd import _lib
    k = _k()
    L, Lt, nrhs = 6, 40, 8
    print(f"[paired x] {prec} rtol={rtol} max_iter={max_iter}", flush=True)
    phi = make_phi(L, Lt, "gauss", seed=41, nchain=nrhs)   # one field per chain: different stopping iterations
    R = make_vec(L, Lt, seed=42, nrhs=nrhs)
    tab = O.precon_table(Lt)
    rd, cd = (torch.float32, torch.complex64) if prec == "f32" else (torch.float64, torch.complex128)
    phd = _t(phi.reshape(nrhs, -1), dev, rd)
    bd = _t(np.stack([O.apply_Mdag(phi[i], R[i], L, Lt) for i in range(nrhs)]), dev, cd)
    tbd = _t(tab, dev, rd)
    out = {}
    for on in (1, 0):
        prev = _lib.lib().qf_set_paired_x_updates(on)
        try:
            if prec == "mixed":
                x, it, err = k.pcg_mixed_solve(phd, bd, L, Lt, ptab=tbd, rtol=rtol, max_iter=max_iter, ru_period=50)
            else:
                x, it, err = k.pcg_solve(phd, bd, L, Lt, ptab=tbd, rtol=rtol, max_iter=max_iter)
            torch.cuda.synchronize()
            out[on] = (x.clone(), it.clone(), err.clone())
        finally:
            _lib.lib().qf_set_paired_x_updates(prev)
    assert _lib.lib().qf_set_paired_x_updates(1) == 1   # the default is restored
    assert torch.equal(out[1][1], out[0][1])
    assert torch.equal(out[1][2], out[0][2])
    assert torch.equal(out[1][0], out[0][0]), (out[1][0] - out[0][0]).abs().max().item()
    its = [int(v) for v in out[1][1].cpu()]
    print(f"[paired x] iterations {its}", flush=True)
    if prec != "mixed" and rtol > 0:
        assert max(its) < max_iter - 1 and {v % 2 for v in its} == {0, 1}, its


@pytest.mark.parametrize("L,Lt,kind,rtol,max_iter", [(6, 60, "uniform", 1e-10, 1200), (8, 40, "gauss", 1e-10, 600),
                                                     (4, 30, "gauss", 1e-13, 200)])
def test_pcg_mixed(dev, L, Lt, kind, rtol, max_iter):
    """Mixed precision with adaptive reliable updates reaches the FP64 true residual; the
    (4, 30) gauss field converges fast, where a fixed long update period lets the FP32
    iterated residual drift and stall around 1e-8."""
    k = _k()
    print(f"[pcg mixed] L={L} Lt={Lt} {kind}", flush=True)
    nrhs = 2
    phi = make_phi(L, Lt, kind, seed=21, nchain=nrhs)
    R = make_vec(L, Lt, seed=22, nrhs=nrhs)
    tab = O.precon_table(Lt)
    b = np.stack([O.apply_Mdag(phi[i], R[i], L, Lt) for i in range(nrhs)])
    x, iters, err = k.pcg_mixed_solve(_t(phi.reshape(nrhs, -1), dev, torch.float64), _t(b, dev, torch.complex128), L, Lt,
                                      ptab=_t(tab, dev, torch.float64), rtol=rtol, max_iter=max_iter, ru_period=50)
    x = x.cpu().numpy()
    for i in range(nrhs):
        true_res = relerr(O.apply_MdagM(phi[i], x[i], L, Lt), b[i])
        assert true_res <= 1.5 * rtol, true_res
        assert abs(err[i].item() - true_res) < 1e-3 * true_res + 1e-15
        xr, _, _ = O.pcg(phi[i], b[i], L, Lt, table=tab, rtol=1e-13, max_iter=4000, early_exit=True)
        assert relerr(x[i], xr) < 1e-7, relerr(x[i], xr)


@pytest.mark.parametrize("L,Lt,nrhs,by", [(2, 4, 1, 0), (4, 10, 2, 0), (6, 12, 1, 6), (12, 24, 2, 4), (36, 6, 1, 0)])
def test_force_f64(dev, L, Lt, nrhs, by):
    k = _k()
    if by and L % by:
        pytest.skip("patch height does not divide L")
    print(f"[force f64] L={L} Lt={Lt} by={by}", flush=True)
    k.set_tuning(1, by, 0, 0)
    try:
        phi = make_phi(L, Lt, "uniform", seed=31, nchain=nrhs)
        xi = make_vec(L, Lt, seed=32, nrhs=nrhs)
        F = k.fermion_force(_t(phi.reshape(nrhs, -1), dev, torch.float64), _t(xi, dev, torch.complex128), L, Lt).cpu().numpy()
        for i in range(nrhs):
            assert relerr(F[i], O.force(phi[i], xi[i], L, Lt)) < 1e-12
    finally:
        k.set_tuning(1, 0, 0, 0)


def test_force_f32(dev):
    k = _k()
    L, Lt = 6, 12
    phi = make_phi(L, Lt, "uniform", seed=33)
    xi = make_vec(L, Lt, seed=34)
    F = k.fermion_force(_t(phi.reshape(1, -1), dev, torch.float32), _t(xi, dev, torch.complex64), L, Lt).cpu().numpy()
    assert relerr(F[0], O.force(phi[0], xi[0], L, Lt)) < 2e-5


@pytest.mark.parametrize("compact", [False, True])
def test_boson_sector(dev, compact):
    k = _k()
    L, Lt, nrhs = 6, 20, 2
    cJ, cK = O.couplings(1.0, 1.0, 0.1)
    phi = make_phi(L, Lt, "gauss", seed=41, nchain=nrhs)
    x = _t(phi.reshape(nrhs, -1), dev, torch.float64)
    Fp, Ft = k.boson_force(x, L, Lt, cJ, cK, compact)
    for i in range(nrhs):
        assert relerr(Fp[i].cpu().numpy(), O.force_b_plaq(phi[i], L, Lt, cK)) < 1e-13
        assert relerr(Ft[i].cpu().numpy(), O.force_b_tau(phi[i], L, Lt, cJ, compact)) < 1e-13
    p = np.random.default_rng(5).standard_normal(phi.shape)
    H, parts = k.hamiltonian(x, L, Lt, cJ, cK, compact, p=_t(p.reshape(nrhs, -1), dev, torch.float64))
    for i in range(nrhs):
        ref = O.action_tau(phi[i], L, Lt, cJ, compact) + O.action_plaq(phi[i], L, Lt, cK) + 0.5 * np.sum(p[i] ** 2)
        assert abs(H[i].item() - ref) < 1e-11 * abs(ref)


def test_leapfrog_boson(dev):
    k = _k()
    L, Lt, nrhs = 4, 10, 2
    cJ, cK = O.couplings(1.0, 1.0, 0.1)
    phi = make_phi(L, Lt, "gauss", seed=51, nchain=nrhs)
    p = np.random.default_rng(52).standard_normal(phi.shape)
    x = _t(phi.reshape(nrhs, -1), dev, torch.float64)
    pt = _t(p.reshape(nrhs, -1), dev, torch.float64)
    Fp, Ft = k.boson_force(x, L, Lt, cJ, cK, False)
    dt = torch.tensor([0.06, 0.03], dtype=torch.float64, device=dev)
    k.leapfrog_boson(x, pt, Fp, Ft, dt, L, Lt, cJ, cK, False, m=1.0, nsub=5)
    for i in range(nrhs):
        xs, ps = phi[i].copy(), p[i].copy()
        h = [0.06, 0.03][i]
        fp = O.force_b_plaq(xs, L, Lt, cK)
        ft = O.force_b_tau(xs, L, Lt, cJ)
        for _ in range(5):
            ps = ps + (fp + ft) * h / 2 / 5
            xs = xs + ps * h / 5
            fp = O.force_b_plaq(xs, L, Lt, cK)
            ft = O.force_b_tau(xs, L, Lt, cJ)
            ps = ps + (fp + ft) * h / 2 / 5
        assert relerr(x[i].cpu().numpy(), xs) < 1e-13
        assert relerr(pt[i].cpu().numpy(), ps) < 1e-13


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-5)])
@pytest.mark.parametrize("shared", [False, True])
def test_operator_from_hop_table(dev, prec, tol, shared):
    """Operators applied from the precomputed hop table equal the oracle."""
    k = _k()
    L, Lt, nrhs = 6, 14, 3
    rd, cd = (torch.float64, torch.complex128) if prec == "f64" else (torch.float32, torch.complex64)
    phi = make_phi(L, Lt, "uniform", seed=91, nchain=1 if shared else nrhs)
    x = make_vec(L, Lt, seed=92, nrhs=nrhs)
    hop = k.hop_table(_t(phi.reshape(phi.shape[0], -1), dev, rd), L, Lt)
    for op in ("M", "Mdag", "MdagM"):
        y = k.apply_hop(op, hop, _t(x, dev, cd), L, Lt).cpu().numpy()
        for r in range(nrhs):
            assert relerr(y[r], OPS[op](phi[0 if shared else r], x[r], L, Lt)) < tol, op


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-5)])
@pytest.mark.parametrize("L,by", [(36, 2), (40, 2), (36, 4)])
def test_operator_repeat_multiwave(dev, prec, tol, L, by):
    """Work-groups of 3 to 7 waves (the whole slice in one work-group) over
    several tau-run lengths, ten launches each: every launch must match the
    oracle.  Guards the exchange barriers and the output stores against
    timing-dependent corruption (a store-data hazard once corrupted the last
    four lanes of each 16-lane group only in some launches)."""
    k = _k()
    Lt = 40
    print(f"[operator repeat] L={L} by={by} {prec}", flush=True)
    phi = make_phi(L, Lt, "uniform", seed=L)
    x = make_vec(L, Lt, seed=L + 1, nrhs=3)
    ref = [O.apply_MdagM(phi[0], x[i], L, Lt) for i in range(3)]
    rd, cd = (torch.float64, torch.complex128) if prec == "f64" else (torch.float32, torch.complex64)
    pi = 1 if prec == "f64" else 0
    pd, xd = _t(phi.reshape(1, -1), dev, rd), _t(x, dev, cd)
    hop = k.hop_table(pd, L, Lt)
    try:
        for T in (1, 2, 3, 5):
            k.set_tuning(pi, by, T, 0)
            for _ in range(10):
                for y in (k.apply("MdagM", pd, xd, L, Lt), k.apply_hop("MdagM", hop, xd, L, Lt)):
                    yn = y.cpu().numpy()
                    for i in range(3):
                        assert relerr(yn[i], ref[i]) < tol, (T, i)
    finally:
        k.set_tuning(pi, 0, 0, 0)


@pytest.mark.parametrize("L,Lt,nrhs,shared,by", [(36, 12, 5, False, 2), (36, 12, 7, True, 2), (12, 20, 9, False, 2),
                                                  (24, 10, 4, False, 4), (6, 12, 11, False, 2), (32, 8, 5, False, 2)])
def test_rhs_per_group(dev, L, Lt, nrhs, shared, by):
    """Plain M^dag M from phi with several right-hand sides side by side in one work-group:
    bit-identical to one per work-group for every group size (the lanes of each slice compute
    the same patches; the last group of a batch that ns does not divide leaves lanes idle), in
    both precisions, per-chain and shared fields; FP64 within 1e-12 of the oracle."""
    import hashlib
    k = _k()
    print(f"[rhs per group] L={L} Lt={Lt} nrhs={nrhs} shared={shared} by={by}", flush=True)
    nch = 1 if shared else nrhs
    phi = make_phi(L, Lt, "uniform", seed=L + nrhs, nchain=nch)
    x = make_vec(L, Lt, seed=L + 2 * nrhs, nrhs=nrhs)
    try:
        for pi, rd, cd in ((0, torch.float32, torch.complex64), (1, torch.float64, torch.complex128)):
            pd, xd = _t(phi.reshape(nch, -1), dev, rd), _t(x, dev, cd)
            k.set_tuning_shape(pi, L, Lt, nrhs, by, 3, 0)
            seen = {}
            for ns in (1, 2, 3, 4, 7):
                k.set_tuning_ns(pi, L, Lt, nrhs, ns)
                eff = k.get_tiling(pi, nrhs, L, Lt)["ns"]
                y = k.apply("MdagM", pd, xd, L, Lt).cpu().numpy()
                seen[(ns, eff)] = hashlib.sha1(y.tobytes()).hexdigest()
                if pi == 1 and ns == 3:
                    for i in range(nrhs):
                        assert relerr(y[i], O.apply_MdagM(phi[0 if shared else i], x[i], L, Lt)) < 1e-12, i
            assert len(set(seen.values())) == 1, (pi, seen)
            assert max(e for _, e in seen) > 1
    finally:
        k.clear_tuning()


def test_rhs_per_group_ignored_where_not_supported(dev):
    """A request for several right-hand sides per work-group leaves the forms without it
    unchanged: band lattices (L=70) and the M^dag M from the hop table give the same bits as
    without the request, and the band result matches the oracle."""
    import hashlib
    k = _k()
    print("[rhs per group ignored]", flush=True)
    try:
        for L, Lt, nrhs in ((70, 4, 3), (12, 8, 4)):
            phi = make_phi(L, Lt, "uniform", seed=L + 11, nchain=nrhs)
            x = make_vec(L, Lt, seed=L + 12, nrhs=nrhs)
            pd, xd = _t(phi.reshape(nrhs, -1), dev, torch.float32), _t(x, dev, torch.complex64)
            hop = k.hop_table(pd, L, Lt)
            out = {}
            for ns in (0, 3):
                k.set_tuning_ns(0, L, Lt, nrhs, ns)
                ys = [k.apply_hop("MdagM", hop, xd, L, Lt).cpu().numpy()]
                if L == 70:
                    ys.append(k.apply("MdagM", pd, xd, L, Lt).cpu().numpy())
                out[ns] = [hashlib.sha1(y.tobytes()).hexdigest() for y in ys]
            assert out[0] == out[3], (L, out)
            if L == 70:
                assert k.get_tiling(0, nrhs, L, Lt)["ns"] == 1
                for i in range(nrhs):
                    assert relerr(ys[1][i], O.apply_MdagM(phi[i], x[i], L, Lt)) < 2e-5
    finally:
        k.clear_tuning()


@pytest.mark.parametrize("T", [12, 15, 16, 24, 30])
def test_long_tau_runs(dev, T):
    """tau runs of 12 to 30 slices, the autotuner's longest candidates, on Ntau = 50 (so that
    every run length but 25 leaves a shorter last run): M^dag M on per-chain fields from phi
    and from the hop table in both precisions against the oracle, the FP64 PCG against the
    oracle's recurrence, and the mixed solver to its tolerance in the oracle's true residual."""
    k = _k()
    L, Lt, nrhs = 6, 50, 2
    print(f"[long tau runs] T={T}", flush=True)
    phi = make_phi(L, Lt, "gauss", seed=T, nchain=nrhs)
    x = make_vec(L, Lt, seed=T + 1, nrhs=nrhs)
    ref = [O.apply_MdagM(phi[i], x[i], L, Lt) for i in range(nrhs)]
    try:
        for pi, rd, cd, tol in ((1, torch.float64, torch.complex128, 1e-12), (0, torch.float32, torch.complex64, 2e-5)):
            k.set_tuning(pi, 0, T, 0)
            pd, xd = _t(phi.reshape(nrhs, -1), dev, rd), _t(x, dev, cd)
            hop = k.hop_table(pd, L, Lt)
            for y in (k.apply("MdagM", pd, xd, L, Lt), k.apply_hop("MdagM", hop, xd, L, Lt)):
                yn = y.cpu().numpy()
                for i in range(nrhs):
                    assert relerr(yn[i], ref[i]) < tol, (pi, i, relerr(yn[i], ref[i]))
        tab = O.precon_table(Lt)
        b = np.stack([O.apply_Mdag(phi[i], x[i], L, Lt) for i in range(nrhs)])
        phd, bd, tbd = _t(phi.reshape(nrhs, -1), dev, torch.float64), _t(b, dev, torch.complex128), _t(tab, dev, torch.float64)
        xs, iters, _ = k.pcg_solve(phd, bd, L, Lt, ptab=tbd, rtol=1e-9, max_iter=80)
        xs = xs.cpu().numpy()
        for i in range(nrhs):
            xr, it, _ = O.pcg(phi[i], b[i], L, Lt, table=tab, rtol=1e-9, max_iter=80, early_exit=False)
            assert relerr(xs[i], xr) < 1e-8, relerr(xs[i], xr)
            assert int(iters[i].item()) == it
        xm, _, em = k.pcg_mixed_solve(phd, bd, L, Lt, ptab=tbd, rtol=1e-10, max_iter=200, ru_period=50)
        xm = xm.cpu().numpy()
        for i in range(nrhs):
            true_res = relerr(O.apply_MdagM(phi[i], xm[i], L, Lt), b[i])
            assert true_res < 2e-10 and em[i].item() < 1e-10, (true_res, em[i].item())
    finally:
        k.set_tuning(1, 0, 0, 0)
        k.set_tuning(0, 0, 0, 0)
