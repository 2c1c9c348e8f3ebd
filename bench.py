"""Benchmark of the HMC fermion hot path on one MI355X.

Workload (default): L=36, Ltau=360 (466,560 sites), J=1, K=1, dtau=0.1, t=1,
N_leapfrog=5 (6 PCG solves per trajectory), delta_t=0.06, the reference's
manual tau-band preconditioner.  A "step" is one HMC trajectory of a batch of
--bs chains, the whole trajectory replayed from one hipGraph (mixed precision:
complex64 PCG with FP64 dots and adaptive FP64 reliable updates, at most 50
iterations apart).  Tilings are picked by the autotuner first.

Headline: every solve runs to rtol 1e-10 (converged chains freeze; the cap is
1000 iterations) on a seeded field with per-bond Gaussian sigma=0.3 around 0,
where PCG converges in ~750 iterations, so the mixed result can be held to the
FP64 one.  Reported next to it:
  * the same-GPU FP64 plain path: complex128 PCG, one chain per trajectory,
    also replayed from a hipGraph (s per chain-trajectory);
  * relative error of the mixed solution / force / S_f against FP64 on the same
    field and pseudofermion, and Delta H and the final field of one trajectory
    run from the same (phi, p, R) in both precisions;
  * reference_budget_uniform_field: phi uniform in [-pi, pi) with the
    reference's fixed 400-iteration budget per solve (PCG does not converge on
    that field), both paths timed the same way;
  * M^dag M (mdagm_kernel, complex64) GB/s from HIP events around back-to-back
    launches on the launch stream, for --roof-nrhs right-hand sides sharing phi
    (working set far above the 256 MB Infinity Cache), vs the 8 TB/s HBM peak;
  * M^dag M as the trajectory applies it (each chain its own field): the
    roofline object;
  * other_configs: the batching configuration (stochastic-estimator solve at
    L=48, Ltau=480 with 128 Z2 noise vectors on one field; the ragged set of 8
    lattices L in {12,18,24,30} x 2 in one launch per kernel) and a 20-step
    mixed-precision trajectory at L=66, Ltau=660, all with the 400-iteration budget;
  * cpu_baseline: the FP64 OpenMP oracle (oracle/qf_oracle.c, a C port of the
    reference's PCG/force path) timed on this host for a bounded number of PCG
    iterations and scaled to one trajectory.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [...]; prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=36)
    ap.add_argument("--Ltau", type=int, default=360)
    ap.add_argument("--bs", type=int, default=8)
    ap.add_argument("--field", choices=["gauss0.3", "uniform"], default="gauss0.3",
                    help="headline field: per-bond Gaussian sigma=0.3 around 0 (PCG converges) or uniform in [-pi,pi)")
    ap.add_argument("--max-iter", type=int, default=1000, help="PCG iteration cap of the headline (converged) solves")
    ap.add_argument("--rtol", type=float, default=1e-10, help="PCG tolerance of the headline solves")
    ap.add_argument("--ref-max-iter", type=int, default=400,
                    help="the reference's fixed PCG budget (max_iter), used on the uniform field and the extras")
    ap.add_argument("--ref-rtol", type=float, default=1e-9, help="the reference's cg_rtol")
    ap.add_argument("--no-ref-budget", action="store_true", help="skip the uniform-field fixed-budget comparison")
    ap.add_argument("--ru-period", type=int, default=50)
    ap.add_argument("--n-leap", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--roof-nrhs", type=int, default=64)
    ap.add_argument("--roof-chains", type=int, default=32,
                    help="chains (own fields) of the roofline M^dag M; 32 at L=36 is a 477 MB working set")
    ap.add_argument("--roof-reps", type=int, default=20)
    ap.add_argument("--cpu-iters", type=int, default=0, help="oracle PCG iterations for cpu_baseline (0 = auto, ~15 s)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="use the built-in tilings instead of the autotuner")
    ap.add_argument("--probe", choices=["per_chain", "shared", "bond_bond", "pcg"], default=None,
                    help="per_chain / shared: only launch a calibration copy and the roofline M^dag M three times each "
                         "(for rocprofv3 --pmc); bond_bond: only the bond-bond estimators of one chain at L=36, "
                         "Ltau=360, Nrv=40, twice (for rocprofv3); pcg: a calibration copy, then 20 iterations of "
                         "the headline mixed PCG (--bs chains, own fields, no chain stops)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batching (L=48, 128 Z2 RHS; ragged set) and L=66 trajectory measurements")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tiling", default=None,
                    help="with --probe pcg: patch height, tau run and update tile as BY,T,TB instead of the autotuner; "
                         "with --probe per_chain / shared: one or more such tilings separated by ';', three launches each")
    return ap.parse_args()


def mdagm_roofline(torch, K, L, Lt, nrhs, reps, dev, use_hop=True):
    """GB/s of the complex64 M^dag M kernel (mdagm_kernel, the operator of every CG iteration).

    Algorithmic bytes per launch: nrhs * (8N read x + 8N write y) + the field once: 8N for phi
    (float32 pair per site) or 16N for the hop table (complex64 pair per site) it is applied from.
    Timed with HIP events on the stream the kernel is launched on (torch's current stream)."""
    N = L * L * Lt
    g = torch.Generator(device="cpu").manual_seed(7)
    phi = ((torch.rand(1, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    x = torch.randn(nrhs, N, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    if use_hop:
        hop = K.hop_table(phi, L, Lt)
        run = lambda: K.apply_hop("MdagM", hop, x, L, Lt, out=y)
        field = 16 * N
    else:
        run = lambda: K.apply("MdagM", phi, x, L, Lt, out=y)
        field = 8 * N
    for _ in range(3):
        run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        run()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    algo = nrhs * 16 * N + field
    return {"kernel": "mdagm_kernel<float> (M^dag M, complex64, one field shared by the right-hand sides, "
                      + ("from the hop table" if use_hop else "from phi") + ")",
            "nrhs": nrhs, "ms_per_launch": ms, "bytes_per_launch": algo, "GBps": algo / (ms * 1e-3) / 1e9,
            "tiling": K.get_tiling(0, nrhs, L, Lt)}


def timed(torch, fn, reps=1):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


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
    hop = K.hop_table(phi, L, Lt)
    t_op = timed(torch, lambda: K.apply_hop("MdagM", hop, eta, L, Lt, out=x), reps=5)
    out["se_L48_128rhs"] = {"solve_s": t_se, "iterations": it, "ms_per_iteration": t_se / it * 1e3,
                            "mdagm_GBps": (nrv * 16 * N + 16 * N) / t_op / 1e9, "mdagm_ms": t_op * 1e3}
    log(f"[bench] SE L=48 128 RHS: {t_se:.3f} s per {it}-iteration solve, M^dag M {out['se_L48_128rhs']['mdagm_GBps']:.0f} GB/s")
    del eta, b, x, hop
    # -- bond-bond estimators of one chain (get_bond_bond_per_b2: T1, T21, T2, T4 over all
    #    C(40, 2) pairs, Tvv), L=36, Ltau=360, Nrv=40 (the reference's Nrv at this size)
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
    t_rop = timed(torch, lambda: K.apply_ragged(plan, phr, br), reps=10)
    out["ragged_8_lattices"] = {"L": Ls, "solve_s": t_rg, "iterations": it, "ms_per_iteration": t_rg / it * 1e3,
                                "mdagm_ms": t_rop * 1e3, "mdagm_GBps": 24 * plan.nsites / t_rop / 1e9}
    log(f"[bench] ragged batch: {t_rg:.3f} s per {it}-iteration solve, one M^dag M launch {t_rop * 1e3:.3f} ms")
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
                                          "max_residual": eng.err.max().item(), "tiling": tune66}
    log(f"[bench] L=66 20-step mixed trajectory: {t66:.3f} s")
    del eng
    return out


def traffic_of(L, Lt, nrhs, kind, tiling):
    """HBM bytes per launch of an M^dag M measurement at this shape and tiling (patch height and
    tau run; the halo re-reads depend on both), as
    measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (each scaled by the
    same counter's reading for a copy of known size) and recorded in
    profiles/mdagm_traffic.json; None if that shape and tiling were not measured."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "mdagm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    for e in rec.get("entries", []):
        if (e.get("L"), e.get("Ltau"), e.get("nrhs"), e.get("kind")) == (L, Lt, nrhs, kind) and \
                (e["tiling"]["by"], e["tiling"]["t"]) == (tiling["by"], tiling["t"]):
            return e.get("traffic_bytes")
    return None


def mdagm_per_chain(torch, K, L, Lt, nchain, reps, dev, tune=True):
    """M^dag M as the trajectory applies it: nchain chains, each with its own
    field (hop table).  Algorithmic bytes per launch: nchain * (8N read x +
    8N write y + 16N read hop table).  Timed with HIP events on the launch
    stream (torch's current stream) around back-to-back launches."""
    N = L * L * Lt
    if tune:
        from qed_fermion_amd import autotune
        autotune.tune(0, L, Lt, nchain, kind="operator", shared_field=False, persist=False)
    g = torch.Generator(device="cpu").manual_seed(8)
    phi = ((torch.rand(nchain, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    x = torch.randn(nchain, N, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    hop = K.hop_table(phi, L, Lt)
    run = lambda: K.apply_hop("MdagM", hop, x, L, Lt, out=y)
    for _ in range(3):
        run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        run()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    algo = nchain * 32 * N
    return {"kernel": "mdagm_kernel<float> (M^dag M, complex64, each chain its own field, from its hop table)",
            "nchains": nchain, "ms_per_launch": ms, "bytes_per_launch": algo, "GBps": algo / (ms * 1e-3) / 1e9,
            "tiling": K.get_tiling(0, nchain, L, Lt)}


def stream_copy_rate(torch, K, dev, reps):
    """Achievable streaming rate: qf_stream_copy of 512 MiB (twice the 256 MB Infinity Cache),
    one 16-byte element per thread, timed like the operator with HIP events on the launch
    stream.  Bytes counted: read + write."""
    nbytes = 512 << 20
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).uniform_()
    b = torch.empty_like(a)
    for _ in range(3):
        K.stream_copy(a, b)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        K.stream_copy(a, b)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    if not torch.equal(a, b):
        raise RuntimeError("stream copy produced a wrong result")
    # the same bytes with one work-group per chunk, each copying its chunk front to back: as many
    # concurrent address streams as the roofline M^dag M's 480-1280 work-groups run
    chunked = {}
    for nch in (480, 960, 1920):
        chunk = (nbytes // nch) // 16 * 16
        b.zero_()
        for _ in range(3):
            K.stream_copy_chunks(a, b, chunk)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            K.stream_copy_chunks(a, b, chunk)
        e1.record(stream)
        torch.cuda.synchronize()
        msc = e0.elapsed_time(e1) / reps
        if not torch.equal(a, b):
            raise RuntimeError("chunked stream copy produced a wrong result")
        chunked[str(nch)] = {"chunk_bytes": chunk, "ms_per_launch": msc, "GBps": 2 * nbytes / (msc * 1e-3) / 1e9}
    return {"kernel": "stream_copy_kernel (one 16-byte element per thread)", "bytes_each_way": nbytes,
            "ms_per_launch": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9,
            "chunked": {"kernel": "chunk_copy_kernel (one work-group per chunk, front to back)", "by_chunks": chunked}}


def probe_pcg(args):
    """Three launches of a plain copy of known size, then one mixed-precision PCG solve of 20
    iterations at rtol 0 (so that no chain stops) on --bs chains of the headline field, tuned as in
    the full run: under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE the copy's counters give the scale
    of the two PCG kernels' (step A, step B held and applying)."""
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
    phi = torch.from_numpy(make_field(np, args.field, np.random.default_rng(args.seed), bs, N)).to(dev)
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
    """Three launches of a plain copy of known size, then three of the roofline
    M^dag M (per-chain fields: --roof-chains chains; shared: --roof-nrhs RHS on
    one field), tuned as in the full run, or three for each tiling of --tiling
    "BY,T,TB;BY,T,TB;..." in that order.  Under rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE the copy's counters give the scale of the kernel's."""
    import torch
    from qed_fermion_amd import autotune
    from qed_fermion_amd import kernels as K
    L, Lt, dev = args.L, args.Ltau, torch.device("cuda:0")
    N = L * L * Lt
    shared = args.probe == "shared"
    n = args.roof_nrhs if shared else args.roof_chains
    tilings = [tuple(int(v) for v in tl.split(",")) for tl in args.tiling.split(";")] if args.tiling else [None]
    if tilings[0] is None and not args.no_tune:
        autotune.tune(0, L, Lt, n, kind="operator", shared_field=shared, persist=False)
    g = torch.Generator(device="cpu").manual_seed(7)
    phi = ((torch.rand(1 if shared else n, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to(dev, torch.float32)
    x = torch.randn(n, N, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    hop = K.hop_table(phi, L, Lt)
    xc = torch.empty(n * N * 2, dtype=torch.float32, device=dev).uniform_()
    yc = torch.empty_like(xc)
    torch.cuda.synchronize()
    for _ in range(3):
        yc.copy_(xc)
    torch.cuda.synchronize()
    order = []
    for tl in tilings:
        if tl is not None:
            K.set_tuning_shape(0, L, Lt, n, *tl)
        for _ in range(3):
            K.apply_hop("MdagM", hop, x, L, Lt, out=y)
        torch.cuda.synchronize()
        order.append(K.get_tiling(0, n, L, Lt))
    print(json.dumps({"probe": args.probe, "L": L, "Ltau": Lt, "n": n, "copy_bytes_each_way": xc.numel() * 4,
                      "algo_read": n * 8 * N + (1 if shared else n) * 16 * N, "algo_write": n * 8 * N,
                      "tilings_in_launch_order": order}), flush=True)


def make_field(np, kind, rng, bs, N):
    """Seeded synthetic boson field [bs, 2N] (kernel layout): 'uniform' in [-pi, pi), or
    'gauss0.3', per-bond Gaussian with sigma = 0.3 around 0."""
    if kind == "uniform":
        return rng.uniform(-np.pi, np.pi, (bs, 2 * N))
    return rng.normal(0.0, 0.3, (bs, 2 * N))


def trajectory_pair(torch, TrajectoryEngine, args, dev, phi0, max_iter, rtol, steps, warmup, tag):
    """Mixed/batched (bs chains) and FP64 plain (1 chain) HMC trajectories, each replayed from one
    hipGraph, from the same field (the plain path takes the first chain's).  Returns timings, PCG
    iterations and residuals."""
    L, Lt, bs = args.L, args.Ltau, args.bs
    common = dict(dtau=0.1, J=1.0, K=1.0, compact=False, m=1.0, delta_t=0.06, n_leap=args.n_leap,
                  max_iter=max_iter, rtol=rtol, device=dev)
    log(f"[bench] {tag}: mixed engine L={L} Ltau={Lt} bs={bs} max_iter={max_iter} rtol={rtol}")
    torch.manual_seed(args.seed)
    eng = TrajectoryEngine(L, Lt, bs, precision="mixed", ru_period=args.ru_period, **common)
    eng.set_phi(phi0)
    t0 = time.time()
    eng.capture(n_warmup=1)
    torch.cuda.synchronize()
    log(f"[bench] {tag}: captured mixed trajectory graph in {time.time() - t0:.1f}s")
    # every trajectory starts from phi0 (a copy of 16N bytes per chain before each replay), so the
    # timed trajectories of both paths see the same fields rather than a chain evolved further
    for _ in range(warmup):
        eng.set_phi(phi0)
        eng.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.set_phi(phi0)
        eng.replay()
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / steps
    out = {"mixed_s_per_batched_trajectory": t_mixed, "mixed_s_per_chain_trajectory": t_mixed / bs,
           "mixed_mean_iters": eng.iters.mean().item(), "mixed_max_residual": eng.err.max().item(),
           "acceptance": eng.accept.float().mean().item()}
    del eng
    log(f"[bench] {tag}: mixed {t_mixed:.3f} s per batched trajectory ({t_mixed / bs:.4f} s per chain), "
        f"mean iters {out['mixed_mean_iters']:.1f}, max residual {out['mixed_max_residual']:.2e}")
    torch.manual_seed(args.seed)
    plain = TrajectoryEngine(L, Lt, 1, precision="fp64", **common)
    plain.set_phi(phi0[:1])
    plain.capture(n_warmup=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        plain.set_phi(phi0[:1])
        plain.replay()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / args.plain_steps
    out.update({"fp64_plain_s_per_trajectory": t_plain, "fp64_mean_iters": plain.iters.mean().item(),
                "fp64_max_residual": plain.err.max().item(), "speedup_vs_fp64_plain": t_plain / (t_mixed / bs)})
    del plain
    log(f"[bench] {tag}: fp64 plain {t_plain:.4f} s per chain-trajectory, mean iters {out['fp64_mean_iters']:.1f}; "
        f"speedup {out['speedup_vs_fp64_plain']:.2f}x")
    return out


def accuracy(torch, K, TrajectoryEngine, args, dev, phi_s, max_iter, rtol):
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
    common = dict(dtau=0.1, J=1.0, K=1.0, compact=False, m=1.0, delta_t=0.06, n_leap=args.n_leap,
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
    rep.update({"dH_fp64": h1 - h0, "dH_abs_err_vs_fp64": abs((m1 - m0) - (h1 - h0)),
                "H0_rel_err_vs_fp64": abs(m0 - h0) / abs(h0),
                "field_rel_err_after_trajectory": (torch.linalg.norm(xmt - x64t) / torch.linalg.norm(x64t)).item()})
    return rep


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

    # ---------------------------------------------------------- tilings (autotuner)
    tilings = {}
    if not args.no_tune:
        from qed_fermion_amd import autotune
        t0 = time.time()
        for prec, nrhs, kind in ((0, bs, "pcg"), (1, 1, "pcg"), (1, bs, "pcg"), (0, args.roof_nrhs, "operator")):
            hit = autotune.tune(prec, L, Lt, nrhs, kind=kind, persist=False)
            tilings[f"{kind}:{'c64' if prec == 0 else 'c128'}:nrhs={nrhs}"] = hit
        log(f"[bench] autotuned in {time.time() - t0:.1f}s: {tilings}")
    rng = np.random.default_rng(args.seed)

    # ---------------------------------------------------------- headline: converged solves
    phi0 = torch.from_numpy(make_field(np, args.field, rng, bs, N)).to(dev)
    head = trajectory_pair(torch, TrajectoryEngine, args, dev, phi0, args.max_iter, args.rtol, args.steps,
                              args.warmup, f"headline ({args.field})")
    acc_report = accuracy(torch, K, TrajectoryEngine, args, dev, phi0[:1].clone(), args.max_iter, args.rtol)
    log(f"[bench] accuracy vs FP64: {acc_report}")

    # ---------------------------------------------------------- the reference's budget on the uniform field
    ref_budget = None
    if not args.no_ref_budget:
        phiu = torch.from_numpy(make_field(np, "uniform", np.random.default_rng(args.seed + 7), bs, N)).to(dev)
        ref_budget = trajectory_pair(torch, TrajectoryEngine, args, dev, phiu, args.ref_max_iter,
                                               args.ref_rtol, 1, 1, "uniform field, fixed budget")
        acc_u = accuracy(torch, K, TrajectoryEngine, args, dev, phiu[:1].clone(), args.ref_max_iter, args.ref_rtol)
        ref_budget.update({"field": "uniform in [-pi,pi)", "max_iter": args.ref_max_iter, "rtol": args.ref_rtol,
                           "accuracy_vs_fp64": acc_u,
                           "note": "PCG with the manual tau-band preconditioner does not converge on this field "
                                   "(FP64 reaches a relative residual of about 6e-6 after 60000 iterations at L=36), "
                                   "so mixed and FP64 iterates differ at any fixed budget; this leg times the "
                                   "reference's fixed 400-iteration trajectory workload"})
        del phiu

    # ---------------------------------------------------------- M^dag M roofline
    roof = mdagm_roofline(torch, K, L, Lt, args.roof_nrhs, args.roof_reps, dev)
    roof_phi = mdagm_roofline(torch, K, L, Lt, args.roof_nrhs, args.roof_reps, dev, use_hop=False)
    roof_b = mdagm_roofline(torch, K, L, Lt, bs, args.roof_reps * 4, dev)
    log(f"[bench] M^dag M nrhs={args.roof_nrhs}: {roof['GBps']:.0f} GB/s (hop table), {roof_phi['GBps']:.0f} GB/s (phi); "
        f"nrhs={bs}: {roof_b['GBps']:.0f} GB/s")

    per_chain = mdagm_per_chain(torch, K, L, Lt, args.roof_chains, args.roof_reps, dev, tune=not args.no_tune)
    log(f"[bench] M^dag M, {args.roof_chains} chains with their own fields: {per_chain['GBps']:.0f} GB/s")
    copy = stream_copy_rate(torch, K, dev, args.roof_reps)
    log(f"[bench] streaming copy reference: {copy['GBps']:.0f} GB/s")
    extra = None if args.no_extras else extras(torch, K, args, dev)

    # ---------------------------------------------------------- CPU baseline (oracle)
    cpu = None
    solves = args.n_leap + 1
    if not args.skip_cpu:
        from oracle import oracle as O
        phi_np = phi0[0].cpu().numpy().reshape(Lt, L, L, 2)
        R = (rng.standard_normal(N) + 1j * rng.standard_normal(N)) / 2.0
        b = O.apply_Mdag(phi_np, R, L, Lt)
        tab = O.precon_table(Lt)
        t0 = time.perf_counter()
        O.pcg(phi_np, b, L, Lt, table=tab, rtol=0.0, max_iter=5, early_exit=False)
        per_it = (time.perf_counter() - t0) / 5
        n_it = args.cpu_iters or max(5, min(2000, int(15.0 / max(per_it, 1e-6))))
        t0 = time.perf_counter()
        O.pcg(phi_np, b, L, Lt, table=tab, rtol=0.0, max_iter=n_it, early_exit=False)
        per_it = (time.perf_counter() - t0) / n_it
        iters = head["fp64_mean_iters"]
        cpu_traj = per_it * iters * solves
        cpu = {"value": cpu_traj, "unit": "s/trajectory", "cores": O.num_threads(), "kind": "port",
               "sample": f"{n_it} FP64 PCG iterations (oracle/qf_oracle.c, OpenMP) at L={L} Ltau={Lt} on the headline "
                         f"field, {per_it * 1e3:.1f} ms/iteration, scaled to {solves} solves x {iters:.0f} iterations "
                         f"(the FP64 path's mean per solve; force and boson steps not included)"}
        log(f"[bench] cpu baseline: {per_it * 1e3:.1f} ms/iter -> {cpu_traj:.1f} s/trajectory")

    value = head["mixed_s_per_chain_trajectory"]
    field_desc = ("per-bond Gaussian sigma=0.3 around 0" if args.field == "gauss0.3" else "uniform in [-pi,pi)")
    out = {
        "metric": "HMC trajectory s + M†M GB/s at L=36,Nτ=360: mixed/batched vs FP64 plain",
        "value": value,
        "unit": "s/trajectory",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["mixed_s_per_batched_trajectory"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64-reliable",
        "data": f"synthetic: phi {field_desc} from a seed, p ~ N(0,1), R ~ randn(complex)/sqrt(2)",
        "config": {"workload": f"HMC trajectory, L={L}, Ltau={Lt}, {bs} chains batched, N_leapfrog={args.n_leap} "
                               f"({solves} PCG solves), every solve to rtol={args.rtol} (cap {args.max_iter}), "
                               f"phi {field_desc}, precon=manual tau band, one hipGraph per trajectory",
                   "L": L, "Ltau": Lt, "chains": bs, "field": args.field, "max_iter": args.max_iter, "rtol": args.rtol,
                   "ru_period": args.ru_period, "parallelism": "single GPU, chains batched"},
        "fp64_plain_s_per_trajectory": head["fp64_plain_s_per_trajectory"],
        "speedup_vs_fp64_plain": head["speedup_vs_fp64_plain"],
        "pcg_iters": {"mixed_mean": head["mixed_mean_iters"], "fp64_mean": head["fp64_mean_iters"]},
        "mixed_max_residual": head["mixed_max_residual"],
        "fp64_max_residual": head["fp64_max_residual"],
        "acceptance": head["acceptance"],
        "accuracy_vs_fp64": acc_report,
        "reference_budget_uniform_field": ref_budget,
        "mdagm_batch_GBps": roof_b["GBps"],
        "mdagm_from_phi_GBps": roof_phi["GBps"],
        "mdagm_per_chain": per_chain,
        "other_configs": extra,
        "roofline": {"bound": "hbm", "achieved": per_chain["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": per_chain["GBps"] / HBM_PEAK_GBS, "traffic": traffic_of(L, Lt, per_chain["nchains"], "per_chain", per_chain["tiling"]),
                     "kernel": per_chain["kernel"], "nchains": per_chain["nchains"],
                     "bytes_per_launch": per_chain["bytes_per_launch"], "ms_per_launch": per_chain["ms_per_launch"],
                     "tiling": per_chain["tiling"], "frac_of_stream_copy": per_chain["GBps"] / copy["GBps"],
                     "frac_of_chunked_copy": per_chain["GBps"] / max(c["GBps"] for c in copy["chunked"]["by_chunks"].values())},
        "stream_copy": copy,
        "mdagm_shared_field": {"GBps": roof["GBps"], "frac": roof["GBps"] / HBM_PEAK_GBS, "kernel": roof["kernel"],
                               "nrhs": roof["nrhs"], "bytes_per_launch": roof["bytes_per_launch"],
                               "ms_per_launch": roof["ms_per_launch"], "tiling": roof["tiling"],
                               "traffic": traffic_of(L, Lt, roof["nrhs"], "shared", roof["tiling"])},
        "tilings": tilings,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
