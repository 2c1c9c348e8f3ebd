"""Benchmark of the HMC fermion hot path on one MI355X.

Headline workload: L=36, Ltau=360 (466,560 sites), J=1, K=1, dtau=0.1, t=1,
N_leapfrog=5 (6 PCG solves per trajectory), the reference's manual tau-band
preconditioner, --bs chains batched.  A "step" is one HMC trajectory of the batch,
replayed from one hipGraph (mixed precision: complex64 PCG with FP64 dots and adaptive
FP64 reliable updates, at most 50 iterations apart; every solve to rtol 1e-10).

The chains start from the reference's initial state (staggered pi flux, the first bs-1
chains plus tau-independent Gaussian noise of widths 0.02..0.08,
hmc_sampler_batch.py:1026-1043) and are thermalised by --therm untimed trajectories that
evolve them (accept/reject) while the reference's per-chain step-size control
(adjust_delta_t, hmc_sampler_batch.py:699-724) brings the acceptance into its window.
The timed trajectories continue the same chains; nothing is reset between them.  The PCG
iteration caps are the reference's gear mechanism (_MAX_ITERS_TO_CAPTURE,
hmc_sampler_batch.py:305-309), one per solve of the trajectory: after the first
thermalisation trajectories each solve's cap is a margin over the iterations it takes,
raised if a solve reaches it (a converged batch still runs its graph to the cap, at about
28 us per empty iteration of dispatch latency).

Reported next to it:
  * the same GPU's FP64 plain path: complex128 PCG, at the same batch and with one chain,
    from the same thermalised state, also replayed from a hipGraph;
  * mixed-precision error against FP64 (solution, force, S_f, Delta H, final field);
  * the roofline kernel: M^dag M applied from phi (the form PCG step A uses), one field per
    chain, --roof-chains chains (working set well above the 256 MB Infinity Cache),
    algorithmic bytes 24 per site per chain (read x and phi, write y), timed with HIP
    events on its launch stream;
  * the same operator from the hop table, the shared-field (stochastic estimator) form,
    and the mixed PCG iteration;
  * extras: the previous headline's sigma=0.3 Gaussian field, the reference's fixed
    400-iteration budget on a uniform field, and the other configurations (batching: SE
    solve at L=48 with 128 Z2 noise vectors, the ragged set; bond-bond estimators;
    L=66 trajectory);
  * cpu_baseline: the FP64 OpenMP oracle (oracle/qf_oracle.c) running one full headline
    trajectory (solves, forces, boson leapfrog) on this host, plus C1 (L=6) trajectories
    and M^dag M rates.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [...]; prints one JSON line.
"""
import argparse
import collections
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "HMC trajectory s + M†M GB/s at L=36,Nτ=360: mixed/batched vs FP64 plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=36)
    ap.add_argument("--Ltau", type=int, default=360)
    ap.add_argument("--bs", type=int, default=8)
    ap.add_argument("--therm", type=int, default=100, help="untimed thermalisation trajectories (dt adapted)")
    ap.add_argument("--delta-t", type=float, default=0.06, help="initial leapfrog step (the reference's delta_t)")
    ap.add_argument("--max-iter", type=int, default=1000, help="PCG iteration cap before the gear is chosen")
    ap.add_argument("--gear-margin", type=float, default=1.1,
                    help="each solve's cap = margin x the largest iteration count it took during thermalisation")
    ap.add_argument("--rtol", type=float, default=1e-10, help="PCG tolerance of the headline solves")
    ap.add_argument("--ref-max-iter", type=int, default=400,
                    help="the reference's fixed PCG budget (max_iter), used on the uniform field and the extras")
    ap.add_argument("--ref-rtol", type=float, default=1e-9, help="the reference's cg_rtol")
    ap.add_argument("--no-ref-budget", action="store_true", help="skip the uniform-field and sigma=0.3 legs")
    ap.add_argument("--ru-period", type=int, default=50)
    ap.add_argument("--n-leap", type=int, default=5)
    ap.add_argument("--fp64-steps", type=int, default=3, help="timed FP64 trajectories (batched and one chain)")
    ap.add_argument("--roof-nrhs", type=int, default=64)
    ap.add_argument("--roof-chains", type=int, default=32,
                    help="chains (own fields) of the roofline M^dag M; 32 at L=36 is a 358 MB working set")
    ap.add_argument("--roof-reps", type=int, default=20)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-c2", action="store_true", help="time the full C2 trajectory (L=24, 40 leapfrog) on the CPU instead of its first 4 steps")
    ap.add_argument("--no-tune", action="store_true", help="use the built-in tilings instead of the autotuner")
    ap.add_argument("--probe", choices=["per_chain", "per_chain_hop", "shared", "bond_bond", "pcg", "persistent"], default=None,
                    help="per_chain (from phi) / per_chain_hop / shared: a calibration copy, then the roofline M^dag M "
                         "three times per tiling (for rocprofv3 --pmc); bond_bond: the bond-bond estimators of one chain "
                         "at L=36, Ltau=360, Nrv=40, twice; pcg: a calibration copy, then 20 iterations of the headline "
                         "mixed PCG (--bs chains, own fields, no chain stops); persistent: the L=48 128-RHS and ragged "
                         "solves as two launches per iteration and as one persistent launch each")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batching (L=48, 128 Z2 RHS; ragged set), bond-bond and L=66 measurements")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tiling", default=None,
                    help="with --probe pcg: patch height, tau run and update tile as BY,T,TB instead of the autotuner; "
                         "with the operator probes: one or more such tilings separated by ';', three launches each, "
                         "each optionally with a fourth field NS (right-hand sides per work-group)")
    return ap.parse_args()


# ------------------------------------------------------------------------------ fields
def pi_flux_start(np, rng, bs, L, Lt):
    """The reference's initial state in the kernels' layout [bs, Ltau*L*L*2]: staggered pi flux
    (x-link sign*pi/4, y-link -sign*pi/4, sign = (-1)^(x+y)) for every chain, the first bs-1
    plus tau-independent Gaussian noise of standard deviations linspace(0.02, 0.08, bs-1)
    (hmc_sampler_batch.py:1026-1043)."""
    x = np.arange(L)[None, :]
    y = np.arange(L)[:, None]
    sign = 1 - 2 * ((x + y) % 2)   # [y][x]
    phi = np.zeros((bs, Lt, L, L, 2))
    phi[..., 0] = sign * math.pi / 4
    phi[..., 1] = -sign * math.pi / 4
    levels = np.linspace(0.02, 0.08, bs - 1) if bs > 1 else []
    for c, s in enumerate(levels):
        phi[c] += s * rng.standard_normal((1, L, L, 2))
    return phi.reshape(bs, -1)


def make_field(np, kind, rng, bs, N):
    """Seeded synthetic boson field [bs, 2N] (kernel layout): 'uniform' in [-pi, pi), or
    'gauss0.3', per-bond Gaussian with sigma = 0.3 around 0."""
    if kind == "uniform":
        return rng.uniform(-np.pi, np.pi, (bs, 2 * N))
    return rng.normal(0.0, 0.3, (bs, 2 * N))


# ------------------------------------------------------------------------------ timing
def events_ms(torch, fn, reps, warm=3):
    """Average ms per call of fn with HIP events on torch's current stream (the launch stream)."""
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def timed(torch, fn, reps=1):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


# ------------------------------------------------------------------------------ roofline
def mdagm_per_chain(torch, K, L, Lt, nchain, reps, dev, source="phi", tune=True):
    """M^dag M with every chain on its own field: from phi (the coefficients' sincos in the
    kernel, as PCG step A applies it) or from the hop table.  Algorithmic bytes per launch:
    nchain * 24N from phi (x read 8N, y write 8N, phi read 8N) or nchain * 32N from the hop
    table (16N of coefficients)."""
    N = L * L * Lt
    if tune:
        from qed_fermion_amd import autotune
        autotune.tune(0, L, Lt, nchain, kind="operator", shared_field=False, source=source, persist=False)
    g = torch.Generator(device="cpu").manual_seed(8)
    phi = ((torch.rand(nchain, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    x = torch.randn(nchain, N, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    if source == "phi":
        run = lambda: K.apply("MdagM", phi, x, L, Lt, out=y)
        per = 24
    else:
        hop = K.hop_table(phi, L, Lt)
        run = lambda: K.apply_hop("MdagM", hop, x, L, Lt, out=y)
        per = 32
    ms = events_ms(torch, run, reps)
    algo = nchain * per * N
    return {"kernel": f"mdagm_kernel<float> (M^dag M, complex64, each chain its own field, from {source})",
            "nchains": nchain, "ms_per_launch": ms, "bytes_per_launch": algo, "bytes_per_site_per_chain": per,
            "GBps": algo / (ms * 1e-3) / 1e9, "frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "tiling": K.get_tiling(0, nchain, L, Lt)}


def mdagm_shared(torch, K, L, Lt, nrhs, reps, dev):
    """M^dag M for nrhs right-hand sides sharing one field (the stochastic estimator's solve),
    from the hop table: nrhs * 16N + 16N algorithmic bytes per launch."""
    N = L * L * Lt
    g = torch.Generator(device="cpu").manual_seed(7)
    phi = ((torch.rand(1, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    x = torch.randn(nrhs, N, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    hop = K.hop_table(phi, L, Lt)
    ms = events_ms(torch, lambda: K.apply_hop("MdagM", hop, x, L, Lt, out=y), reps)
    algo = nrhs * 16 * N + 16 * N
    return {"kernel": "mdagm_kernel<float> (M^dag M, complex64, one field shared by the right-hand sides, hop table)",
            "nrhs": nrhs, "ms_per_launch": ms, "bytes_per_launch": algo, "GBps": algo / (ms * 1e-3) / 1e9,
            "frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "tiling": K.get_tiling(0, nrhs, L, Lt)}


def pcg_iteration_rate(torch, K, L, Lt, bs, dev, ru_period, iters=60):
    """Mixed PCG at the headline batch with rtol 0 (no chain stops): ms per iteration and the
    rate on the CG-iteration budget of 96 bytes per site per chain (M^dag M 24, P 16, the
    fused vector updates and dots 56)."""
    N = L * L * Lt
    g = torch.Generator(device="cpu").manual_seed(9)
    phi = ((torch.rand(bs, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev)
    b = torch.randn(bs, N, generator=g, dtype=torch.complex128).to(dev)
    ptab = K.precon_table(Lt, device=dev)
    x = torch.empty_like(b)
    phi32 = phi.float()
    ms = events_ms(torch, lambda: K.pcg_mixed_solve(phi, b, L, Lt, ptab=ptab, rtol=0.0, max_iter=iters, ru_period=ru_period,
                                                    phi32=phi32, x=x), 3, warm=1) / iters
    algo = bs * 96 * N
    return {"chains": bs, "ms_per_iteration": ms, "bytes_per_iteration": algo, "GBps": algo / (ms * 1e-3) / 1e9,
            "frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "note": "mixed PCG iteration including the reliable-update checks; 96 B/site/chain CG-iteration budget"}


def traffic_of(L, Lt, nrhs, kind, tiling):
    """HBM bytes per launch of an M^dag M measurement at this shape and tiling (patch height and
    tau run; the halo re-reads depend on both), as measured by rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE (each scaled by the same counter's reading for a copy of known size) and
    recorded in profiles/mdagm_traffic.json; None if that shape and tiling were not measured."""
    path = os.path.join(ROOT, "profiles", "mdagm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    for e in rec.get("entries", []):
        if (e.get("L"), e.get("Ltau"), e.get("nrhs"), e.get("kind")) == (L, Lt, nrhs, kind) and \
                (e["tiling"]["by"], e["tiling"]["t"], e["tiling"].get("ns", 1)) == (tiling["by"], tiling["t"], tiling.get("ns", 1)):
            return e.get("traffic_bytes")
    return None


def copy_reference(torch, dev, reps):
    """torch's device copy of 512 MiB (twice the Infinity Cache), HIP events: a plain streaming
    rate on this box, read + write bytes."""
    nbytes = 512 << 20
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).uniform_()
    b = torch.empty_like(a)
    ms = events_ms(torch, lambda: b.copy_(a), reps)
    return {"kernel": "torch copy_ (512 MiB)", "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}


# ------------------------------------------------------------------------------ chains
class Chains:
    """One TrajectoryEngine replayed from its hipGraph, with the chains evolving from one
    trajectory to the next, device-side logs of Delta H and acceptance, and the PCG gear."""

    def __init__(self, torch, TrajectoryEngine, args, dev, bs, precision, phi0, dt, max_iter, tag):
        self.torch, self.tag, self.bs = torch, tag, bs
        common = dict(dtau=0.1, J=1.0, K=1.0, compact=False, m=1.0, delta_t=args.delta_t, n_leap=args.n_leap,
                      rtol=args.rtol, device=dev)
        self.eng = TrajectoryEngine(args.L, args.Ltau, bs, precision=precision, ru_period=args.ru_period,
                                    max_iter=max_iter, **common)
        self.eng.set_phi(phi0)
        self.eng.dt.copy_(dt)
        self.capture()

    def capture(self):
        """(Re)capture the trajectory graph with the current caps, keeping the chains' state."""
        x, dt = self.eng.x.clone(), self.eng.dt.clone()
        self.eng.graph = None
        t0 = time.time()
        self.eng.capture(n_warmup=1)
        self.torch.cuda.synchronize()
        self.eng.x.copy_(x)
        self.eng.dt.copy_(dt)
        log(f"[bench] {self.tag}: captured trajectory graph (caps {self.eng.max_iters}) in {time.time() - t0:.1f}s")

    def set_caps(self, caps):
        if list(caps) != self.eng.max_iters:
            self.eng.max_iter = list(caps)
            self.capture()

    def thermalise(self, n, adapt, gear_margin, cap_max, first=5):
        """n evolving trajectories; the reference's dt control per chain (adapt) and the PCG
        gears: after `first` trajectories each solve of the trajectory gets its own cap, margin x
        the most iterations it has taken, raised by 20% when a solve reaches it, and set once
        more from everything seen at the end."""
        from qed_fermion_amd.trajectory import adjust_delta_t
        eng = self.eng
        queues = [collections.deque(maxlen=5) for _ in range(self.bs)]
        seen = [0.0] * (eng.n_leap + 1)
        acc, dH = [], []
        gear = lambda: [min(cap_max, 10 * math.ceil(gear_margin * max(v, 1) / 10)) for v in seen]
        for k in range(n):
            eng.replay()
            th = eng.acceptance().tolist()
            seen = [max(a, b) for a, b in zip(seen, eng.iters.max(dim=1).values.tolist())]
            acc.append(eng.accept.float().mean().item())
            dH.append((eng.H1 - eng.H0).mean().item())
            if adapt:
                for b in range(self.bs):
                    queues[b].append(th[b])
                adjust_delta_t(eng.dt, queues)
            hit = (eng.status == 1).any(dim=1).tolist()
            if k + 1 == first:
                self.set_caps(gear())
            elif k + 1 > first and any(hit):
                self.set_caps([min(cap_max, max(g, int(1.2 * c))) if h else c
                               for g, c, h in zip(gear(), eng.max_iters, hit)])
        self.set_caps(gear())
        return {"trajectories": n, "final_dt": eng.dt.tolist(), "acceptance_last20": sum(acc[-20:]) / max(1, len(acc[-20:])),
                "mean_dH_last20": sum(dH[-20:]) / max(1, len(dH[-20:])), "max_iters_seen_per_solve": seen,
                "caps_per_solve": list(eng.max_iters)}

    def timed(self, steps, warmup):
        torch, eng = self.torch, self.eng
        dH = torch.zeros(steps, self.bs, dtype=torch.float64, device=eng.device)
        acc = torch.zeros(steps, self.bs, dtype=torch.float64, device=eng.device)
        its = torch.zeros(steps, dtype=torch.float64, device=eng.device)
        maxit = torch.zeros(steps, self.bs, dtype=torch.int32, device=eng.device)
        for _ in range(warmup):
            eng.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            eng.replay()
            torch.sub(eng.H1, eng.H0, out=dH[k])
            acc[k].copy_(eng.accept)
            its[k] = eng.iters.mean()
            maxit[k] = (eng.status == 1).sum(dim=0, dtype=torch.int32)
        torch.cuda.synchronize()
        s = (time.perf_counter() - t0) / steps
        d = dH.cpu()
        return {"s_per_batched_trajectory": s, "s_per_chain_trajectory": s / self.bs, "steps": steps, "warmup": warmup,
                "acceptance": acc.mean().item(), "mean_dH": d.mean().item(), "mean_exp_minus_dH": (-d).exp().mean().item(),
                "mean_pcg_iters": its.mean().item(), "solves_at_cap": int(maxit.sum().item()),
                "max_residual": eng.err.max().item(), "caps_per_solve": list(eng.max_iters), "dt": eng.dt.tolist()}


# ------------------------------------------------------------------------------ accuracy
def accuracy(torch, K, TrajectoryEngine, args, dev, phi_s, dt, max_iter, rtol):
    """Mixed precision against FP64 on one state: one solve (solution, force, S_f) and one eager
    trajectory from the same (phi, p, R) in both precisions (Delta H, final field)."""
    L, Lt = args.L, args.Ltau
    N = L * L * Lt
    g = torch.Generator(device="cpu").manual_seed(args.seed + 1)
    R = (torch.randn(1, N, generator=g, dtype=torch.complex128) / math.sqrt(2)).to(dev)
    p = torch.randn(1, 2 * N, generator=g, dtype=torch.float64).to(dev)
    psi = K.apply("Mdag", phi_s, R, L, Lt)
    ptab64 = K.precon_table(Lt, dtype=torch.float64, device=dev)
    x64, it64, e64 = K.pcg_solve(phi_s, psi, L, Lt, 0.1, ptab=ptab64, rtol=rtol, max_iter=max_iter)
    xm, itm, em = K.pcg_mixed_solve(phi_s, psi, L, Lt, 0.1, ptab=ptab64, rtol=rtol, max_iter=max_iter,
                                    ru_period=args.ru_period)
    F64 = K.fermion_force(phi_s, x64, L, Lt)
    Fm = K.fermion_force(phi_s, xm, L, Lt)
    Sf64 = torch.vdot(psi.view(-1), x64.view(-1)).real.item()
    Sfm = torch.vdot(psi.view(-1), xm.view(-1)).real.item()
    rep = {"solution_rel_err_vs_fp64": (torch.linalg.norm(xm - x64) / torch.linalg.norm(x64)).item(),
           "force_rel_err_vs_fp64": (torch.linalg.norm(Fm - F64) / torch.linalg.norm(F64)).item(),
           "Sf_rel_err_vs_fp64": abs(Sfm - Sf64) / abs(Sf64),
           "residual_fp64": e64.item(), "residual_mixed_true": em.item(),
           "iters_fp64": it64.item(), "iters_mixed": itm.item()}
    common = dict(dtau=0.1, J=1.0, K=1.0, compact=False, m=1.0, delta_t=float(dt), n_leap=args.n_leap,
                  max_iter=max_iter, rtol=rtol, device=dev)
    res = {}
    for prec in ("fp64", "mixed"):
        e = TrajectoryEngine(L, Lt, 1, precision=prec, ru_period=args.ru_period, **common)
        e.set_phi(phi_s)
        e.p.copy_(p)
        e.R.copy_(R)
        e.propose(draw=False)
        torch.cuda.synchronize()
        res[prec] = (e.H0.item(), e.H1.item(), e.x.clone())
        del e
    (h0, h1, x64t), (m0, m1, xmt) = res["fp64"], res["mixed"]
    rep.update({"dt": float(dt), "dH_fp64": h1 - h0, "dH_abs_err_vs_fp64": abs((m1 - m0) - (h1 - h0)),
                "H0_rel_err_vs_fp64": abs(m0 - h0) / abs(h0),
                "field_rel_err_after_trajectory": (torch.linalg.norm(xmt - x64t) / torch.linalg.norm(x64t)).item()})
    return rep


# ------------------------------------------------------------------------------ extras
def reset_leg(torch, TrajectoryEngine, args, dev, phi0, max_iter, rtol, tag):
    """Mixed/batched and FP64 one-chain trajectories, each replayed once from the same field
    (reset before every replay): the earlier headline's form, kept as an extra."""
    out = {}
    for prec, bs in (("mixed", args.bs), ("fp64", 1)):
        eng = TrajectoryEngine(args.L, args.Ltau, bs, precision=prec, ru_period=args.ru_period, dtau=0.1, J=1.0, K=1.0,
                               compact=False, m=1.0, delta_t=0.06, n_leap=args.n_leap, max_iter=max_iter, rtol=rtol,
                               device=dev)
        eng.set_phi(phi0[:bs])
        eng.capture(n_warmup=1)
        eng.set_phi(phi0[:bs])
        eng.replay()
        torch.cuda.synchronize()
        eng.set_phi(phi0[:bs])
        t0 = time.perf_counter()
        eng.replay()
        torch.cuda.synchronize()
        s = time.perf_counter() - t0
        out[prec] = {"s_per_chain_trajectory": s / bs, "chains": bs, "mean_iters": eng.iters.mean().item(),
                     "max_residual": eng.err.max().item(), "dH_mean": (eng.H1 - eng.H0).mean().item()}
        del eng
    out["speedup"] = out["fp64"]["s_per_chain_trajectory"] / out["mixed"]["s_per_chain_trajectory"]
    log(f"[bench] {tag}: mixed {out['mixed']['s_per_chain_trajectory']:.4f} s/chain-traj, fp64 "
        f"{out['fp64']['s_per_chain_trajectory']:.4f}, {out['speedup']:.2f}x")
    return out


def extras(torch, K, args, dev):
    """The other configurations, one measurement each (not the headline)."""
    import numpy as np
    from qed_fermion_amd import autotune
    from qed_fermion_amd.trajectory import TrajectoryEngine
    out = {}
    # -- batching: stochastic-estimator solve, L=48, Ltau=480, 128 Z2 noise vectors, one field
    L, Lt, nrv = 48, 480, 128
    N = L * L * Lt
    g = torch.Generator(device="cpu").manual_seed(48)
    phi = ((torch.rand(1, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    eta = (torch.randint(0, 2, (nrv, N), generator=g) * 2 - 1).to(torch.float32)
    eta = (torch.complex(eta, (torch.randint(0, 2, (nrv, N), generator=g) * 2 - 1).to(torch.float32)) / math.sqrt(2)).to(dev)
    if not args.no_tune:
        autotune.tune(0, L, Lt, nrv, kind="operator", persist=False)
    b = K.apply("Mdag", phi, eta, L, Lt)
    ptab = K.precon_table(Lt, dtype=torch.float32, device=dev)
    x = torch.empty_like(b)
    it = args.ref_max_iter
    t_se = timed(torch, lambda: K.pcg_solve(phi, b, L, Lt, ptab=ptab, rtol=0.0, max_iter=it, x=x))
    t_sep = timed(torch, lambda: K.pcg_solve_persistent(phi, b, L, Lt, ptab=ptab, rtol=0.0, max_iter=it, x=x))
    hop = K.hop_table(phi, L, Lt)
    t_op = timed(torch, lambda: K.apply_hop("MdagM", hop, eta, L, Lt, out=x), reps=5)
    out["se_L48_128rhs"] = {"solve_s": t_se, "iterations": it, "ms_per_iteration": t_se / it * 1e3,
                            "mdagm_GBps": (nrv * 16 * N + 16 * N) / t_op / 1e9, "mdagm_ms": t_op * 1e3,
                            "persistent_solve_s": t_sep,
                            "note": "solve_s: two launches per iteration; persistent_solve_s: the whole solve as one launch"}
    log(f"[bench] SE L=48 128 RHS: {t_se:.3f} s per {it}-iteration solve, M^dag M {out['se_L48_128rhs']['mdagm_GBps']:.0f} GB/s")
    del eta, b, x, hop
    # -- bond-bond estimators of one chain, L=36, Ltau=360, Nrv=40 (the reference's Nrv at this size)
    from qed_fermion_amd.observables import BondBondEstimator
    L, Lt, nrv = 36, 360, 40
    N = L * L * Lt
    z2 = lambda: (torch.randint(0, 2, (nrv, N), generator=g) * 2 - 1).to(torch.float32)
    eta_b = (torch.complex(z2(), z2()) / math.sqrt(2)).to(dev)
    G_b = torch.randn(nrv, N, generator=g, dtype=torch.complex64).to(dev)
    bos = ((torch.rand(1, 2, L, L, Lt, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    est = BondBondEstimator(L, Lt, nrv, device=dev)
    t_bb = timed(torch, lambda: est.bond_bond(eta_b, G_b, bos), reps=3)
    out["bond_bond_L36_nrv40"] = {"s_per_chain": t_bb, "pairs": nrv * (nrv - 1) // 2, "pair_chunk": est.chunk,
                                  "estimators": ["T1", "T21", "T2", "T4", "Tvv"],
                                  "note": "get_bond_bond_per_b2 after the G_eta solve; G_eta synthetic (the cost does not depend on it)"}
    log(f"[bench] bond-bond estimators L=36 Nrv=40: {t_bb * 1e3:.1f} ms per chain")
    del eta_b, G_b, est
    # -- batching: ragged set, L in {12, 18, 24, 30} x 2, Ltau = 10 L, one launch per kernel
    Ls = [12, 18, 24, 30, 12, 18, 24, 30]
    plan = K.RaggedPlan(Ls, [10 * v for v in Ls], device=dev)
    phr = ((torch.rand(2 * plan.nsites, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    br = torch.randn(plan.nsites, generator=g, dtype=torch.complex128).to(dev, torch.complex64)
    ptr_ = plan.precon_table(dtype=torch.float32, device=dev)
    t_rg = timed(torch, lambda: K.pcg_solve_ragged(plan, phr, br, ptab=ptr_, rtol=0.0, max_iter=it))
    t_rgp = timed(torch, lambda: K.pcg_solve_ragged_persistent(plan, phr, br, ptab=ptr_, rtol=0.0, max_iter=it))
    t_rop = timed(torch, lambda: K.apply_ragged(plan, phr, br), reps=10)
    phr64 = phr.double()
    br64 = br.to(torch.complex128)
    t_rgm = timed(torch, lambda: K.pcg_mixed_solve_ragged(plan, phr64, br64, ptab=ptr_, rtol=0.0, max_iter=it,
                                                          ru_period=args.ru_period, phi32=phr))
    out["ragged_8_lattices"] = {"L": Ls, "solve_s": t_rg, "iterations": it, "ms_per_iteration": t_rg / it * 1e3,
                                "mdagm_ms": t_rop * 1e3, "mdagm_GBps": 24 * plan.nsites / t_rop / 1e9,
                                "mixed_solve_s": t_rgm, "mixed_ms_per_iteration": t_rgm / it * 1e3,
                                "persistent_solve_s": t_rgp}
    log(f"[bench] ragged batch: {t_rg:.3f} s per {it}-iteration solve ({t_rgm:.3f} s mixed), one M^dag M launch "
        f"{t_rop * 1e3:.3f} ms")
    # -- differentiation at full size: L=66, Ltau=660, a 20-step mixed trajectory
    L, Lt = 66, 660
    N = L * L * Lt
    rng = np.random.default_rng(66)
    ph66 = torch.from_numpy(rng.uniform(-np.pi, np.pi, (1, 2 * N))).to(dev)
    tune66 = None
    if not args.no_tune:   # one chain of a large lattice: the tau-run length decides how many CUs get work
        tune66 = autotune.tune(0, L, Lt, 1, kind="pcg", persist=False)
        autotune.tune(1, L, Lt, 1, kind="pcg", persist=False)
    eng = TrajectoryEngine(L, Lt, 1, precision="mixed", ru_period=args.ru_period, dtau=0.1, J=1.0, K=1.0, compact=False,
                           m=1.0, delta_t=0.06, n_leap=20, max_iter=args.ref_max_iter, rtol=args.ref_rtol, device=dev)
    eng.set_phi(ph66)
    eng.capture(n_warmup=1)
    t66 = timed(torch, eng.replay)
    out["trajectory_L66_20step_mixed"] = {"s_per_trajectory": t66, "solves": 21, "max_iter": args.ref_max_iter,
                                          "ms_per_pcg_iteration": t66 / (21 * args.ref_max_iter) * 1e3,
                                          "max_residual": eng.err.max().item(), "tiling": tune66}
    log(f"[bench] L=66 20-step mixed trajectory: {t66:.3f} s")
    del eng
    # -- the plain-method configurations: C1 (L=6, Ltau=60, uniform field, 20 leapfrog steps, Metropolis) and
    #    C2 (L=24, Ltau=240, sigma=0.3 field, 40 leapfrog steps), one FP64 chain, every solve to 1e-10, one
    #    hipGraph per trajectory with caps per solve from an eager trajectory on the same field
    for name, (Lc, Ltc, kind, nl, dtc) in (("c1_L6_fp64_20step", (6, 60, "uniform", 20, 0.06)),
                                            ("c2_L24_fp64_40step", (24, 240, "gauss", 40, 0.02))):
        Nc = Lc * Lc * Ltc
        phc = rng.uniform(-np.pi, np.pi, 2 * Nc) if kind == "uniform" else 0.3 * rng.standard_normal(2 * Nc)
        phc = torch.from_numpy(phc).view(1, -1).to(dev)
        tune_c = autotune.tune(1, Lc, Ltc, 1, kind="pcg", persist=False)
        eng = TrajectoryEngine(Lc, Ltc, 1, precision="fp64", dtau=0.1, J=1.0, K=1.0, compact=False, m=1.0, delta_t=dtc,
                               n_leap=nl, max_iter=4000, rtol=1e-10, device=dev)
        eng.set_phi(phc)
        eng.propose()
        torch.cuda.synchronize()
        need = eng.iters[:, 0].tolist()
        eng.max_iter = [min(4000, int(1.2 * v) + 20) for v in need]
        graph = sum(eng.max_iters) <= eng.MAX_GRAPH_ITERATIONS   # else eager launches
        eng.set_phi(phc)
        if graph:
            eng.capture(n_warmup=1)

        def one():
            eng.set_phi(phc)
            eng.replay() if graph else eng.trajectory()
        tc = timed(torch, one, reps=3)
        out[name] = {"s_per_trajectory": tc, "solves": nl + 1, "dt": dtc, "field": kind, "graph": graph, "tiling": tune_c,
                     "iterations_per_solve_eager": [min(need), max(need)], "caps_total": sum(eng.max_iters),
                     "max_residual": eng.err.max().item(), "solves_at_cap": int((eng.status == 1).sum().item())}
        log(f"[bench] {name}: {tc * 1e3:.1f} ms per trajectory")
        del eng
    return out


# ------------------------------------------------------------------------------ CPU
def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(np, O, args, phi_h, dt, cap, rng):
    """The FP64 OpenMP oracle on this host: one full headline trajectory (6 solves to rtol with
    the same cap and early exit, forces, the boson leapfrog) on a thermalised chain, C1
    trajectories, C2's first 4 leapfrog steps (its full trajectory with --cpu-c2), and M^dag M
    rates (48 bytes per site in complex128)."""
    L, Lt = args.L, args.Ltau
    N = L * L * Lt
    tab = O.precon_table(Lt)
    p = rng.standard_normal(2 * N)
    R = (rng.standard_normal(N) + 1j * rng.standard_normal(N)) / 2.0
    t0 = time.perf_counter()
    _, _, H0, H1, err = O.trajectory(phi_h.reshape(Lt, L, L, 2), p.reshape(Lt, L, L, 2), R, L, Lt, dt=dt,
                                     n_leap=args.n_leap, table=tab, rtol=args.rtol, max_iter=cap)
    t_head = time.perf_counter() - t0
    out = {"value": t_head, "unit": "s/trajectory", "cores": O.num_threads(), "kind": "port",
           "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "sample": f"one full FP64 trajectory of the headline workload on one thermalised chain (L={L}, Ltau={Lt}, "
                     f"{args.n_leap} leapfrog steps at dt={dt:.4f}, every solve to rtol={args.rtol} with cap {cap}, "
                     f"forces and the boson leapfrog included), oracle/qf_oracle.c with OpenMP",
           "headline_dH": H1 - H0, "headline_last_residual": err}
    # C1: L=6, Ltau=60, uniform field, 20 leapfrog steps, rtol 1e-10
    L1, Lt1 = 6, 60
    N1 = L1 * L1 * Lt1
    ph1 = rng.uniform(-np.pi, np.pi, (Lt1, L1, L1, 2))
    reps, t0 = 5, time.perf_counter()
    for _ in range(reps):
        R1 = (rng.standard_normal(N1) + 1j * rng.standard_normal(N1)) / 2.0
        O.trajectory(ph1, rng.standard_normal(ph1.shape), R1, L1, Lt1, dt=0.06, n_leap=20, table=O.precon_table(Lt1),
                     rtol=1e-10, max_iter=4000)
    out["c1_s_per_trajectory"] = (time.perf_counter() - t0) / reps
    # C2: L=24, Ltau=240, sigma=0.3 field, dt 0.02: the full 40-step trajectory with --cpu-c2, otherwise a
    # bounded sample of its first 4 leapfrog steps (5 of its 41 solves)
    L2, Lt2 = 24, 240
    N2 = L2 * L2 * Lt2
    ph2 = rng.normal(0.0, 0.3, (Lt2, L2, L2, 2))
    R2 = (rng.standard_normal(N2) + 1j * rng.standard_normal(N2)) / 2.0
    nl2 = 40 if args.cpu_c2 else 4
    t0 = time.perf_counter()
    O.trajectory(ph2, rng.standard_normal(ph2.shape), R2, L2, Lt2, dt=0.02, n_leap=nl2, table=O.precon_table(Lt2),
                 rtol=1e-10, max_iter=4000)
    t2 = time.perf_counter() - t0
    if args.cpu_c2:
        out["c2_s_per_trajectory"] = t2
    else:
        out["c2_sample"] = {"leapfrog_steps": nl2, "solves": nl2 + 1, "s": t2,
                            "s_per_trajectory_scaled_by_solves": t2 * 41 / (nl2 + 1)}
    rates = {}
    for (Lc, Ltc) in ((6, 60), (24, 240), (36, 360)):
        Nc = Lc * Lc * Ltc
        ph = rng.uniform(-np.pi, np.pi, (Ltc, Lc, Lc, 2))
        xv = rng.standard_normal(Nc) + 1j * rng.standard_normal(Nc)
        O.apply_MdagM(ph, xv, Lc, Ltc)
        n = max(3, int(2e7 // Nc))
        t0 = time.perf_counter()
        for _ in range(n):
            O.apply_MdagM(ph, xv, Lc, Ltc)
        s = (time.perf_counter() - t0) / n
        rates[f"L{Lc}_Ltau{Ltc}"] = {"ms": s * 1e3, "GBps": 48 * Nc / s / 1e9}
    out["mdagm_c128"] = rates
    log(f"[bench] cpu baseline: {t_head:.1f} s per headline trajectory on {out['cores']} threads ({out['cpu_model']}), "
        f"C1 {out['c1_s_per_trajectory']:.3f} s, C2 {nl2} steps {t2:.1f} s")
    return out


# ------------------------------------------------------------------------------ probes
def probe_pcg(args):
    """Three launches of a plain copy of known size, then one mixed-precision PCG solve of 20
    iterations at rtol 0 (so that no chain stops) on --bs chains of the thermalised-like
    pi-flux start, tuned as in the full run: under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE the
    copy's counters give the scale of the PCG kernels' (step A, step B held and applying)."""
    import numpy as np
    import torch
    from qed_fermion_amd import autotune
    from qed_fermion_amd import kernels as K
    L, Lt, bs, dev = args.L, args.Ltau, args.bs, torch.device("cuda:0")
    N = L * L * Lt
    if args.tiling:
        by, t, tb = (int(v) for v in args.tiling.split(","))
        K.set_tuning_shape(0, L, Lt, bs, by, t, tb)
    elif not args.no_tune:
        autotune.tune(0, L, Lt, bs, kind="pcg", persist=False)
    phi = torch.from_numpy(pi_flux_start(np, np.random.default_rng(args.seed), bs, L, Lt)).to(dev)
    g = torch.Generator(device="cpu").manual_seed(9)
    b = torch.randn(bs, N, generator=g, dtype=torch.complex128).to(dev)
    ptab = K.precon_table(Lt, device=dev)
    xc = torch.empty(bs * N * 2, dtype=torch.float32, device=dev).uniform_()
    yc = torch.empty_like(xc)
    torch.cuda.synchronize()
    for _ in range(3):
        yc.copy_(xc)
    torch.cuda.synchronize()
    K.pcg_mixed_solve(phi, b, L, Lt, ptab=ptab, rtol=0.0, max_iter=20, ru_period=args.ru_period)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "pcg", "L": L, "Ltau": Lt, "chains": bs, "iterations": 20,
                      "copy_bytes_each_way": xc.numel() * 4, "tiling": K.get_tiling(0, bs, L, Lt)}), flush=True)


def probe_persistent(args):
    """The batching configuration's solves, each as two launches per iteration and as one
    persistent launch, 400 iterations at rtol 0 (no chain stops): 128 complex64 Z2 noise vectors
    at L=48, Ltau=480 on one field, and the ragged set L in {12, 18, 24, 30} x 2.  Under
    rocprofv3 --kernel-trace each persistent solve is one dispatch."""
    import torch
    from qed_fermion_amd import kernels as K
    dev = torch.device("cuda:0")
    it = 400
    L, Lt, nrv = 48, 480, 128
    N = L * L * Lt
    g = torch.Generator(device="cpu").manual_seed(48)
    phi = ((torch.rand(1, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    eta = (torch.randint(0, 2, (nrv, N), generator=g) * 2 - 1).to(torch.float32)
    eta = (torch.complex(eta, (torch.randint(0, 2, (nrv, N), generator=g) * 2 - 1).to(torch.float32)) / math.sqrt(2)).to(dev)
    b = K.apply("Mdag", phi, eta, L, Lt)
    ptab = K.precon_table(Lt, dtype=torch.float32, device=dev)
    x = torch.empty_like(b)
    t_two = timed(torch, lambda: K.pcg_solve(phi, b, L, Lt, ptab=ptab, rtol=0.0, max_iter=it, x=x))
    t_per = timed(torch, lambda: K.pcg_solve_persistent(phi, b, L, Lt, ptab=ptab, rtol=0.0, max_iter=it, x=x))
    del eta, b, x
    Ls = [12, 18, 24, 30, 12, 18, 24, 30]
    plan = K.RaggedPlan(Ls, [10 * v for v in Ls], device=dev)
    phr = ((torch.rand(2 * plan.nsites, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    br = torch.randn(plan.nsites, generator=g, dtype=torch.complex128).to(dev, torch.complex64)
    ptr_ = plan.precon_table(dtype=torch.float32, device=dev)
    r_two = timed(torch, lambda: K.pcg_solve_ragged(plan, phr, br, ptab=ptr_, rtol=0.0, max_iter=it))
    r_per = timed(torch, lambda: K.pcg_solve_ragged_persistent(plan, phr, br, ptab=ptr_, rtol=0.0, max_iter=it))
    print(json.dumps({"probe": "persistent", "iterations": it,
                      "se_L48_128rhs_s": {"two_launch": t_two, "persistent": t_per},
                      "ragged_8_lattices_s": {"two_launch": r_two, "persistent": r_per}}), flush=True)


def probe_bond_bond():
    """The bond-bond estimators of one chain (L=36, Ltau=360, Nrv=40, complex64, synthetic
    G_eta), evaluated twice; prints the time of the second evaluation."""
    import torch
    from qed_fermion_amd.observables import BondBondEstimator
    L, Lt, nrv, dev = 36, 360, 40, torch.device("cuda:0")
    N = L * L * Lt
    g = torch.Generator(device="cpu").manual_seed(1)
    z2 = lambda: (torch.randint(0, 2, (nrv, N), generator=g) * 2 - 1).to(torch.float32)
    eta = (torch.complex(z2(), z2()) / math.sqrt(2)).to(dev)
    G = torch.randn(nrv, N, generator=g, dtype=torch.complex64).to(dev)
    bos = ((torch.rand(1, 2, L, L, Lt, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    est = BondBondEstimator(L, Lt, nrv, device=dev)
    t = timed(torch, lambda: est.bond_bond(eta, G, bos))
    print(json.dumps({"probe": "bond_bond", "L": L, "Ltau": Lt, "nrv": nrv, "method": est.method, "nwg": est.nwg,
                      "ms_per_chain": t * 1e3}), flush=True)


def probe(args):
    """Three launches of a plain copy of known size, then three of the roofline M^dag M
    (per_chain: --roof-chains chains on their own fields, from phi; per_chain_hop: the same from
    the hop table; shared: --roof-nrhs RHS on one field), tuned as in the full run, or three for
    each tiling of --tiling "BY,T,TB;BY,T,TB;..." in that order.  Under rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE the copy's counters give the scale of the kernel's."""
    import torch
    from qed_fermion_amd import autotune
    from qed_fermion_amd import kernels as K
    L, Lt, dev = args.L, args.Ltau, torch.device("cuda:0")
    N = L * L * Lt
    shared = args.probe == "shared"
    source = "phi" if args.probe == "per_chain" else "hop"
    n = args.roof_nrhs if shared else args.roof_chains
    tilings = [tuple(int(v) for v in tl.split(",")) for tl in args.tiling.split(";")] if args.tiling else [None]
    if tilings[0] is None and not args.no_tune:
        autotune.tune(0, L, Lt, n, kind="operator", shared_field=shared, source=source, persist=False)
    g = torch.Generator(device="cpu").manual_seed(7)
    phi = ((torch.rand(1 if shared else n, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    x = torch.randn(n, N, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    hop = K.hop_table(phi, L, Lt) if source == "hop" else None
    run = (lambda: K.apply("MdagM", phi, x, L, Lt, out=y)) if source == "phi" else \
        (lambda: K.apply_hop("MdagM", hop, x, L, Lt, out=y))
    xc = torch.empty(n * N * 2, dtype=torch.float32, device=dev).uniform_()
    yc = torch.empty_like(xc)
    torch.cuda.synchronize()
    for _ in range(3):
        yc.copy_(xc)
    torch.cuda.synchronize()
    order = []
    for tl in tilings:
        if tl is not None:
            K.set_tuning_shape(0, L, Lt, n, *tl[:3])
            K.set_tuning_ns(0, L, Lt, n, tl[3] if len(tl) > 3 else 0)
        for _ in range(3):
            run()
        torch.cuda.synchronize()
        order.append(K.get_tiling(0, n, L, Lt))
    field = (1 if shared else n) * (8 if source == "phi" else 16) * N
    print(json.dumps({"probe": args.probe, "L": L, "Ltau": Lt, "n": n, "copy_bytes_each_way": xc.numel() * 4,
                      "algo_read": n * 8 * N + field, "algo_write": n * 8 * N,
                      "tilings_in_launch_order": order}), flush=True)


# ------------------------------------------------------------------------------ main
def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "qed_fermion_amd runs on one GPU only (--gpus must be 1): chains, pseudofermions and "
                                   "noise vectors would shard across GPUs without communication, but no multi-GPU "
                                   "path is built"}))
        sys.exit(2)
    if args.probe == "bond_bond":
        probe_bond_bond()
        return
    if args.probe == "pcg":
        probe_pcg(args)
        return
    if args.probe == "persistent":
        probe_persistent(args)
        return
    if args.probe:
        probe(args)
        return
    import numpy as np
    import torch

    from qed_fermion_amd import kernels as K
    from qed_fermion_amd.trajectory import TrajectoryEngine

    dev = torch.device("cuda:0")
    L, Lt, bs = args.L, args.Ltau, args.bs
    N = L * L * Lt
    solves = args.n_leap + 1

    # ---------------------------------------------------------- tilings (autotuner)
    tilings = {}
    if not args.no_tune:
        from qed_fermion_amd import autotune
        t0 = time.time()
        for prec, nrhs in ((0, bs), (1, 1), (1, bs)):
            tilings[f"pcg:{'c64' if prec == 0 else 'c128'}:nrhs={nrhs}"] = autotune.tune(prec, L, Lt, nrhs, kind="pcg",
                                                                                           persist=False)
        log(f"[bench] autotuned in {time.time() - t0:.1f}s: {tilings}")
    rng = np.random.default_rng(args.seed)

    # ---------------------------------------------------------- headline: thermalised, evolving chains
    phi_start = torch.from_numpy(pi_flux_start(np, rng, bs, L, Lt)).to(dev)
    dt0 = torch.full((bs,), args.delta_t, dtype=torch.float64, device=dev)
    torch.manual_seed(args.seed)
    mixed = Chains(torch, TrajectoryEngine, args, dev, bs, "mixed", phi_start, dt0, args.max_iter, "mixed/batched")
    t0 = time.time()
    therm = mixed.thermalise(args.therm, adapt=True, gear_margin=args.gear_margin, cap_max=args.max_iter)
    log(f"[bench] thermalised {args.therm} trajectories in {time.time() - t0:.1f}s: {therm}")
    state = mixed.eng.x.clone()
    dt = mixed.eng.dt.clone()
    cap = list(mixed.eng.max_iters)
    head = mixed.timed(args.steps, args.warmup)
    log(f"[bench] mixed/batched: {head['s_per_chain_trajectory']:.4f} s per chain-trajectory, acceptance "
        f"{head['acceptance']:.2f}, <dH> {head['mean_dH']:.3f}, iters {head['mean_pcg_iters']:.0f}")
    del mixed
    torch.manual_seed(args.seed + 11)
    f64b = Chains(torch, TrajectoryEngine, args, dev, bs, "fp64", state, dt, cap, "fp64/batched")
    head64b = f64b.timed(args.fp64_steps, 1)
    del f64b
    torch.manual_seed(args.seed + 12)
    f641 = Chains(torch, TrajectoryEngine, args, dev, 1, "fp64", state[:1], dt[:1], cap, "fp64/one chain")
    head641 = f641.timed(args.fp64_steps, 1)
    del f641
    log(f"[bench] fp64 batched {head64b['s_per_chain_trajectory']:.4f}, fp64 one chain "
        f"{head641['s_per_chain_trajectory']:.4f} s per chain-trajectory")
    acc_report = accuracy(torch, K, TrajectoryEngine, args, dev, state[:1].clone(), dt[0].item(), max(cap), args.rtol)
    log(f"[bench] accuracy vs FP64: {acc_report}")

    # ---------------------------------------------------------- extra legs (reset fields)
    legs = None
    if not args.no_ref_budget:
        legs = {"gauss0.3_to_1e-10": reset_leg(torch, TrajectoryEngine, args, dev,
                                              torch.from_numpy(make_field(np, "gauss0.3", np.random.default_rng(args.seed), bs, N)).to(dev),
                                              1000, 1e-10, "sigma=0.3 field, reset"),
                "uniform_fixed_400": reset_leg(torch, TrajectoryEngine, args, dev,
                                               torch.from_numpy(make_field(np, "uniform", np.random.default_rng(args.seed + 7), bs, N)).to(dev),
                                               args.ref_max_iter, args.ref_rtol, "uniform field, 400 fixed")}
        legs["note"] = ("fields reset before every replay, not equilibrium HMC: the sigma=0.3 Gaussian relaxes (large "
                        "negative dH) and PCG does not converge on the uniform field (the reference's fixed budget)")

    # ---------------------------------------------------------- M^dag M roofline
    per_chain = mdagm_per_chain(torch, K, L, Lt, args.roof_chains, args.roof_reps, dev, "phi", tune=not args.no_tune)
    per_chain_hop = mdagm_per_chain(torch, K, L, Lt, args.roof_chains, args.roof_reps, dev, "hop", tune=not args.no_tune)
    if not args.no_tune:
        from qed_fermion_amd import autotune
        autotune.tune(0, L, Lt, args.roof_nrhs, kind="operator", persist=False)
    shared = mdagm_shared(torch, K, L, Lt, args.roof_nrhs, args.roof_reps, dev)
    pcg_it = pcg_iteration_rate(torch, K, L, Lt, bs, dev, args.ru_period)
    copy = copy_reference(torch, dev, args.roof_reps)
    log(f"[bench] M^dag M per chain from phi: {per_chain['GBps']:.0f} GB/s ({per_chain['frac']:.3f} of peak); "
        f"hop {per_chain_hop['GBps']:.0f}; shared {shared['GBps']:.0f}; PCG iteration {pcg_it['GBps']:.0f}; copy {copy['GBps']:.0f}")
    extra = None if args.no_extras else extras(torch, K, args, dev)

    # ---------------------------------------------------------- CPU baseline (oracle)
    cpu = None
    if not args.skip_cpu:
        from oracle import oracle as O
        cpu = cpu_baseline(np, O, args, state[0].cpu().numpy(), dt[0].item(), max(cap), rng)

    value = head["s_per_chain_trajectory"]
    t64_1 = head641["s_per_chain_trajectory"]
    t64_b = head64b["s_per_chain_trajectory"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "s/trajectory",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["s_per_batched_trajectory"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64-reliable",
        "data": "synthetic: the reference's pi-flux + noise start, thermalised by HMC; p ~ N(0,1), R ~ randn(complex)/sqrt(2)",
        "config": {"workload": f"HMC trajectory, L={L}, Ltau={Lt}, {bs} chains batched, N_leapfrog={args.n_leap} "
                               f"({solves} PCG solves), every solve to rtol={args.rtol} (caps per solve {cap}), chains thermalised "
                               f"from pi flux + noise ({args.therm} trajectories, dt adapted) and evolving, precon=manual "
                               f"tau band, one hipGraph per trajectory",
                   "L": L, "Ltau": Lt, "chains": bs, "field": "pi_flux_noise_thermalised", "rtol": args.rtol,
                   "max_iter_per_solve": cap, "ru_period": args.ru_period, "parallelism": "single GPU, chains batched"},
        "headline": head,
        "thermalisation": therm,
        "fp64_batched_s_per_trajectory": t64_b,
        "fp64_plain_s_per_trajectory": t64_1,
        "fp64_batched": head64b,
        "fp64_plain": head641,
        "speedup_vs_fp64_plain": t64_1 / value,
        "speedup_split": {"batching_fp64": t64_1 / t64_b, "precision_batched": t64_b / value},
        "accuracy_vs_fp64": acc_report,
        "reset_field_legs": legs,
        "roofline": {"bound": "hbm", "achieved": per_chain["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": per_chain["frac"],
                     "traffic": traffic_of(L, Lt, per_chain["nchains"], "per_chain_phi", per_chain["tiling"]),
                     "kernel": per_chain["kernel"], "nchains": per_chain["nchains"],
                     "bytes_per_launch": per_chain["bytes_per_launch"], "bytes_per_site_per_chain": 24,
                     "ms_per_launch": per_chain["ms_per_launch"], "tiling": per_chain["tiling"],
                     "frac_of_copy_reference": per_chain["GBps"] / copy["GBps"]},
        "mdagm_per_chain_hop": per_chain_hop,
        "mdagm_shared_field": dict(shared, traffic=traffic_of(L, Lt, shared["nrhs"], "shared", shared["tiling"])),
        "pcg_iteration": pcg_it,
        "copy_reference": copy,
        "other_configs": extra,
        "tilings": tilings,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
