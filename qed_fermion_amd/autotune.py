"""Autotuner for the work-group tiling of the fermion kernels.

Three shape parameters decide how a launch maps onto the MI355X:

* ``by``  -- patch height of the operator kernels (each thread owns a 2 x by
  patch of a tau slice; one work-group owns a whole slice), in {2, 4, 6}; on a
  lattice too wide for a whole slice of 2 x 2 patches, by = 2 means bands of
  rows with recomputed halos (several work-groups per slice);
* ``t``   -- tau slices marched by one operator work-group (the tau pipeline
  pays one extra B application per run, so long runs are cheaper, but fewer
  work-groups fill fewer CUs);
* ``tb``  -- tau slices of a PCG update tile (16, 32, 64 or 128; the tile's
  13-slice preconditioner window re-reads 12 halo slices);
* ``ns``  -- right-hand sides per work-group of the plain M^dag M from phi: several
  slices side by side fill whole wavefronts where one slice does not (L=36: 324 lanes
  in six wavefronts, 648 in eleven).

The reference picks a fixed CUDA block (BLOCK_SIZE = (4, 8),
hmc_sampler_batch.py:111); here the best values depend on L, Ltau, the batch
and the precision, so they are measured: :func:`tune_operator` times
M^dag M applications, :func:`tune_pcg` times a fixed number of PCG
iterations (both kernels of the iteration).  Winners are registered with the
library per shape (``qf_set_tuning_shape``) and kept in a JSON cache keyed by
the device, so later processes start from :func:`load`.
"""
import json
import math
import os

import torch

from . import kernels
from ._lib import QF_PREC_F32, QF_PREC_F64

_CACHE_ENV = "QF_TUNING_CACHE"


def cache_path():
    return os.environ.get(_CACHE_ENV, os.path.join(os.path.expanduser("~"), ".cache", "qed_fermion_amd", "tuning.json"))


def device_key(dev=None):
    p = torch.cuda.get_device_properties(dev if dev is not None else torch.cuda.current_device())
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|{p.multi_processor_count}"


def _shape_key(prec, L, Lt, nrhs, kind):
    return f"{kind}:{prec}:{L}:{Lt}:{nrhs}"


def valid_by(L, prec):
    """Patch heights that fit L (2 x by patches tile the slice; 512-thread
    work-groups, 1024 for by = 2).  complex128 with by >= 4 exceeds the
    register file, so it is only offered when nothing else fits.  Where a whole
    slice of 2 x 2 patches does not fit one work-group, by = 2 is offered as
    bands (the only choice when no other height fits)."""
    out = []
    for by in (2, 4, 6):
        if L % by or by % 2:
            continue
        nt = (L // 2) * (L // by)
        if nt > (1024 if by == 2 else 512):
            continue
        out.append(by)
    if prec == QF_PREC_F64 and 2 in out:
        out = [2]
    if 2 not in out and L % 2 == 0 and 2 * (1024 // (L // 2)) >= 18:
        out.append(2)   # bands of 2 x 2 patches with 8 halo rows (L <= 226)
    return out


def ns_candidates(L, by, nrhs):
    """Right-hand sides per work-group to time for the plain M^dag M from phi: 1 up to what
    one work-group holds (1024 threads with 2 x 2 patches, 512 above) and the batch."""
    if by <= 0 or L % by or nrhs <= 1:
        return [1]
    nts = (L // 2) * (L // by)
    top = min(nrhs, (1024 if by == 2 else 512) // nts, 4)
    return list(range(1, max(top, 1) + 1))


def _t_candidates(Lt, nrhs=None, ncu=None):
    """tau-run lengths to time: a fixed ladder, plus for nrhs right-hand sides (work-group
    columns) the shortest runs whose work-groups (nrhs * ceil(Lt / t)) still fit k per CU for
    k = 1..4, so that the last round of work-groups is not a partial one (32 chains at
    Ltau = 360: t = 23 fills 512 work-groups, where t = 24 leaves 480)."""
    out = {t for t in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 19, 22, 24) if t <= Lt}
    if nrhs:
        if ncu is None:
            ncu = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
        for k in (1, 2, 3, 4):
            runs = (k * ncu) // nrhs
            if runs >= 1:
                t = -(-Lt // runs)
                if t <= min(Lt, 48):
                    out.add(t)
    return sorted(out)


def _time(fn, reps, warm=2):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _refine(table, select, fn, reps, scale=1.0, top=3, rounds=3):
    """Re-time the ``top`` fastest entries of ``table`` in ``rounds`` interleaved rounds and keep
    each one's minimum (launch-to-launch noise only adds time), so that two tilings a few
    percent apart are not ranked by one noisy sample."""
    best = sorted(table, key=table.get)[:top]
    for _ in range(rounds):
        for k in best:
            select(k)
            table[k] = min(table[k], _time(fn, reps) * scale)


def _dtypes(prec):
    return (torch.float32, torch.complex64) if prec == QF_PREC_F32 else (torch.float64, torch.complex128)


def tune_operator(prec, L, Lt, nrhs, shared_field=True, reps=10, seed=0, source="hop"):
    """Best (by, t, ns) for M^dag M on nrhs right-hand sides, applied from the hop table
    (source "hop") or from phi (source "phi", the coefficients' sincos in the kernel; ns
    right-hand sides per work-group are timed only there); one field shared by all RHS if
    shared_field, else one per RHS.  Returns (by, t, ms_per_launch, table, ns)."""
    rd, cd = _dtypes(prec)
    g = torch.Generator(device="cpu").manual_seed(seed)
    N = L * L * Lt
    nchain = 1 if shared_field else nrhs
    phi = ((torch.rand(nchain, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to("cuda", rd)
    x = torch.randn(nrhs, N, generator=g, dtype=torch.complex128).to("cuda", cd)
    y = torch.empty_like(x)
    if source == "hop":
        hop = kernels.hop_table(phi, L, Lt)
        run = lambda: kernels.apply_hop("MdagM", hop, x, L, Lt, out=y)
    elif source == "phi":
        run = lambda: kernels.apply("MdagM", phi, x, L, Lt, out=y)
    else:
        raise ValueError(f"source {source!r}")
    def select(k):
        kernels.set_tuning_shape(prec, L, Lt, nrhs, k[0], k[1], 0)
        kernels.set_tuning_ns(prec, L, Lt, nrhs, k[2])

    table = {}
    for by in valid_by(L, prec):
        for ns in (ns_candidates(L, by, nrhs) if source == "phi" else [1]):
            for t in _t_candidates(Lt, -(-nrhs // ns)):
                select((by, t, ns))
                table[(by, t, ns)] = _time(run, reps)
    _refine(table, select, run, reps)
    select((0, 0, 0))
    (by, t, ns), ms = min(table.items(), key=lambda kv: kv[1])
    return by, t, ms, table, ns


def tune_pcg(prec, L, Lt, nrhs, iters=20, reps=3, seed=0, tb_list=(16, 32, 64, 128)):
    """Best (by, t, tb) for the batched PCG iteration (per-chain fields):
    times ``iters`` iterations of pcg_solve with rtol = 0 (no chain stops).
    Returns (by, t, tb, ms_per_iteration, table)."""
    rd, cd = _dtypes(prec)
    g = torch.Generator(device="cpu").manual_seed(seed)
    N = L * L * Lt
    phi = ((torch.rand(nrhs, 2 * N, generator=g, dtype=torch.float64) * 2 - 1) * math.pi).to("cuda", rd)
    b = torch.randn(nrhs, N, generator=g, dtype=torch.complex128).to("cuda", cd)
    ptab = kernels.precon_table(Lt, device="cuda", dtype=rd) if Lt >= 26 else None
    x = torch.empty_like(b)

    def run():   # the workspace cache grows when a tiling needs more partial sums
        kernels.pcg_solve(phi, b, L, Lt, ptab=ptab, rtol=0.0, max_iter=iters, x=x)

    table = {}
    for by in valid_by(L, prec):
        for t in _t_candidates(Lt, nrhs):
            kernels.set_tuning_shape(prec, L, Lt, nrhs, by, t, 0)
            table[(by, t, 0)] = _time(run, reps, warm=1) / iters
    (by, t, _), _ = min(table.items(), key=lambda kv: kv[1])
    for tb in tb_list:
        kernels.set_tuning_shape(prec, L, Lt, nrhs, by, t, tb)
        table[(by, t, tb)] = _time(run, reps, warm=1) / iters
    _refine(table, lambda k: kernels.set_tuning_shape(prec, L, Lt, nrhs, *k), run, reps, scale=1.0 / iters)
    kernels.set_tuning_shape(prec, L, Lt, nrhs, 0, 0, 0)
    (by, t, tb), ms = min(table.items(), key=lambda kv: kv[1])
    return by, t, tb, ms, table


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(path, data):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(prec, L, Lt, nrhs, kind="pcg", persist=True, force=False, **kw):
    """Tune one shape and register the result with the library.

    kind = "pcg": the PCG iteration with one field per RHS (the sampler's
    solves); kind = "operator": M^dag M with one field shared by the RHS (the
    stochastic estimator's solves; shared_field=False and source="phi" time it
    with one field per RHS from phi, as PCG step A applies it).  A cached result for this device is
    reused unless force.  Returns {"by", "t", "tb", "ms"}."""
    path = cache_path()
    cache = _load_cache(path) if persist else {}
    dk = device_key()
    key = _shape_key(prec, L, Lt, nrhs, kind)
    if kind == "operator" and (not kw.get("shared_field", True) or kw.get("source", "hop") != "hop"):
        key += f":{'shared' if kw.get('shared_field', True) else 'per_chain'}:{kw.get('source', 'hop')}"
    hit = cache.get(dk, {}).get(key)
    if hit is None or force:
        if kind == "pcg":
            by, t, tb, ms, _ = tune_pcg(prec, L, Lt, nrhs, **kw)
        elif kind == "operator":
            by, t, ms, _, ns = tune_operator(prec, L, Lt, nrhs, **kw)
            tb = 0
        else:
            raise ValueError(f"kind {kind!r}")
        hit = {"by": by, "t": t, "tb": tb, "ms": ms}
        if kind == "operator":
            hit["ns"] = ns
        if persist:
            cache = _load_cache(path)
            cache.setdefault(dk, {})[key] = hit
            _save_cache(path, cache)
    kernels.set_tuning_shape(prec, L, Lt, nrhs, hit["by"], hit["t"], hit["tb"])
    kernels.set_tuning_ns(prec, L, Lt, nrhs, hit.get("ns", 0))
    return dict(hit)


def load(path=None):
    """Register every cached entry of this device with the library; returns
    the number of entries applied."""
    cache = _load_cache(path or cache_path()).get(device_key(), {})
    n = 0
    for key, v in cache.items():
        _kind, prec, L, Lt, nrhs = key.split(":")[:5]
        kernels.set_tuning_shape(int(prec), int(L), int(Lt), int(nrhs), v["by"], v["t"], v["tb"])
        kernels.set_tuning_ns(int(prec), int(L), int(Lt), int(nrhs), v.get("ns", 0))
        n += 1
    return n
