"""On-disk formats of the reference's HMC runs, read and written without executing file contents.

Three kinds of file, all ``torch.save`` of plain dicts of CPU tensors:

* boson ensemble  ``check_points/boson_ensemble/boson_{specifics}_{step}.pt`` holding
  ``{'boson': [bs, 2, Lx, Ly, Ltau], 'delta_t_tensor': [bs]}`` (hmc_sampler_batch.py:3164-3181).
  A bare boson tensor is accepted on load, as the reference accepts it (994-1003).
* run checkpoint  ``check_points/hmc_check_point_{suffix}/ckpt_N_{specifics}_step_{step}.pt``
  holding ``{'boson', 'step', 'cur_step', ...per-sweep lists}`` (3146-3162, 3184-3200).
* P cache         ``preconditioners/precon_ckpt_L_{Lx}_Ltau_{Ltau}_dtau_{dtau}_t_{t}.pt`` holding
  ``{'indices': int64 [2, nnz], 'values': complex64 [nnz], 'size': torch.Size}``
  (preconditioners_orig/precon_manual.py:177-184; read by reset_precon, hmc_sampler_batch.py:461-547).

Every load goes through ``torch.load(..., weights_only=True)``: tensors, numbers,
strings, lists, dicts and ``torch.Size`` load; anything that would need code
from the file raises.  Paths are relative to a caller-given root (the
reference's root is its own package directory).
"""
import os

import torch

from . import precon


def _load(path):
    return torch.load(path, map_location="cpu", weights_only=True)


def default_nrv(Lx, Ltau):
    """Noise vectors of the stochastic estimator by lattice size (hmc_sampler_batch.py:58-62):
    40 for 80000 <= sites <= 1250000, 30 up to 2160000 sites (and below 80000), 20 above."""
    n = Lx * Lx * Ltau
    if 80000 <= n <= 1250000:
        return 40
    if n <= 2160000:
        return 30
    return 20


def specifics(s, gear0_steps=None, Nrv=None, compact=None):
    """The run tag of a sampler (HmcSampler.initialize_specifics, hmc_sampler_batch.py:330-331).

    ``s`` needs Lx, Ltau, N_step, bs, J, K, dtau, Nf, N_leapfrog, m, cg_rtol,
    max_iter, max_tau_block_idx and threshold_queue.  gear0_steps, Nrv and
    compact default to the reference's environment variables (``gear0_steps``
    1000, ``Nrv`` by lattice size, ``compact`` 0; hmc_sampler_batch.py:30,
    58-62, 79) when not given; ``s.compact`` is used for compact when the
    sampler has it.  The reference sizes Nrv from its module-level lattice
    (the ``L``/``asym`` variables), which is the sampler's lattice when both
    come from the environment."""
    if gear0_steps is None:
        gear0_steps = int(os.getenv("gear0_steps", "1000"))
    if Nrv is None:
        Nrv = int(os.getenv("Nrv", str(default_nrv(s.Lx, s.Ltau))))
    if compact is None:
        compact = bool(getattr(s, "compact", int(os.getenv("compact", "0")) != 0))
    Jtau = s.J * s.dtau / s.Nf * 4
    Kin = s.K / s.dtau / s.Nf * 2
    return (f"hmc_{s.Lx}_Ltau_{s.Ltau}_Nstp_{s.N_step}_bs{s.bs}_Jtau_{Jtau:.2g}_K_{Kin:.2g}_dtau_{s.dtau:.2g}"
            f"_delta_{0.028:.2g}_N_leapfrog_{s.N_leapfrog}_m_{s.m:.2g}_cg_rtol_{s.cg_rtol:.2g}"
            f"_max_iter_{s.max_iter}_max_block_idx_{s.max_tau_block_idx}_gear0_steps_{gear0_steps}"
            f"_dt_deque_max_len_{s.threshold_queue[0].maxlen}_Nrv_{Nrv}_cmp_{compact}")


def ensemble_file_name(s, step, spec=None):
    """File name of the ensemble member saved at ``step`` (hmc_sampler_batch.py:3166-3176, 981-991).

    The reference maps every run onto the tag of its production ensembles:
    cg_rtol 1e-15 -> 1e-09, no max_iter, Nstp -> 10000 (6800 at Lx = 60),
    cmp_False, and Jtau -> 1 at Lx = 50 and 60."""
    spec = spec if spec is not None else specifics(s)
    name = f"boson_{spec}_{step}"
    name = name.replace("cg_rtol_1e-15", "cg_rtol_1e-09")
    name = name.replace(f"_max_iter_{s.max_iter}", "")
    name = name.replace(f"Nstp_{s.N_step}", "Nstp_10000")
    name = name.replace("cmp_True", "cmp_False")
    if s.Lx in (50, 60):
        name = name.replace(f"Jtau_{s.J * s.dtau / s.Nf * 4:.2g}", "Jtau_1")
    if s.Lx == 60:
        name = name.replace("Nstp_10000", "Nstp_6800")
    return name + ".pt"


ENSEMBLE_STEPS = tuple(range(4000, 10001, 500))


def save_to_file(res, data_folder, filename):
    """HmcSampler.save_to_file (hmc_sampler_batch.py:3211-3215): ``{data_folder}/{filename}.pt``."""
    os.makedirs(data_folder, exist_ok=True)
    path = os.path.join(data_folder, f"{filename}.pt")
    torch.save(res, path)
    return path


def save_boson_ensemble(s, root, step):
    """Write the sampler's boson and step sizes as the ensemble member of sweep ``step``
    (hmc_sampler_batch.py:3164-3181, where step is the index of the sweep just finished)."""
    folder = os.path.join(root, "check_points", "boson_ensemble")
    name = ensemble_file_name(s, step)[:-3]
    return save_to_file({"boson": s.boson.detach().cpu(), "delta_t_tensor": s.delta_t_tensor.detach().cpu()},
                        folder, name)


def load_boson_ensemble(s, root, steps=ENSEMBLE_STEPS):
    """All ensemble members of the sampler's tag under ``root``: (bosons, delta_t_tensors), CPU tensors."""
    folder = os.path.join(root, "check_points", "boson_ensemble")
    bosons, dts = [], []
    for step in steps:
        path = os.path.join(folder, ensemble_file_name(s, step))
        if not os.path.exists(path):
            continue
        obj = _load(path)
        if isinstance(obj, dict):
            bosons.append(obj["boson"])
            dts.append(obj["delta_t_tensor"])
        elif isinstance(obj, torch.Tensor):
            bosons.append(obj)
        else:
            raise ValueError(f"{path}: neither a boson tensor nor a {{'boson', 'delta_t_tensor'}} dict")
    return bosons, dts


def _check_boson(s, b, what):
    want = (s.bs, 2, s.Lx, s.Ly, s.Ltau)
    if tuple(b.shape) != want:
        raise ValueError(f"{what}: boson shape {tuple(b.shape)}, sampler needs {want}")


def initialize_boson_ensemble(s, root, steps=ENSEMBLE_STEPS):
    """HmcSampler.initialize_boson_ensemble (hmc_sampler_batch.py:971-1016).

    Picks one saved boson (and, independently, one saved delta_t_tensor) at
    random with torch's global generator; without files falls back to
    ``initialize_boson_pi_flux_randn_matfree``.  Returns the number of
    members found.  The boson is converted to the sampler's dtype."""
    bosons, dts = load_boson_ensemble(s, root, steps)
    if not bosons:
        s.initialize_boson_pi_flux_randn_matfree()
        return 0
    b = bosons[torch.randint(len(bosons), (1,)).item()]
    _check_boson(s, b, "ensemble")
    s.boson = b.to(s.device, s.dtype)
    if dts:
        dt = dts[torch.randint(len(dts), (1,)).item()]
        s.delta_t_tensor = dt.to(s.device, s.delta_t_tensor.dtype).reshape(s.bs)
    return len(bosons)


def checkpoint_file_name(s, step, spec=None):
    """``ckpt_N_{specifics}_step_{step}`` (hmc_sampler_batch.py:3161, 3199), without the suffix."""
    return f"ckpt_N_{spec if spec is not None else specifics(s)}_step_{step}"


def save_checkpoint(s, root, suffix="bench", lists=None, step=None):
    """A run checkpoint in the reference's dict layout (hmc_sampler_batch.py:3184-3200).

    ``lists`` adds the per-sweep records (G_list, S_plaq_list, ...; tensors),
    which this package does not compute.  ``step`` names the file and defaults
    to the sampler's sweep counter ``s.step``, which is stored under 'step'."""
    step = s.step if step is None else step
    res = {"boson": s.boson.detach().cpu(), "step": int(s.step), "cur_step": int(s.cur_step),
           "delta_t_tensor": s.delta_t_tensor.detach().cpu()}
    for k, v in (lists or {}).items():
        res[k] = v.detach().cpu() if isinstance(v, torch.Tensor) else v
    folder = os.path.join(root, "check_points", f"hmc_check_point_{suffix}")
    return save_to_file(res, folder, checkpoint_file_name(s, step))


def load_checkpoint(path, s=None):
    """Read a run checkpoint; with a sampler, resume it from the file (boson, step, cur_step
    and, when present, delta_t_tensor).  Returns the dict."""
    res = _load(path)
    if not isinstance(res, dict) or "boson" not in res:
        raise ValueError(f"{path}: not a checkpoint dict with a 'boson' entry")
    if s is not None:
        _check_boson(s, res["boson"], path)
        s.boson = res["boson"].to(s.device, s.dtype)
        s.step = int(res.get("step", 0))
        s.cur_step = int(res.get("cur_step", s.step))
        if "delta_t_tensor" in res:
            s.delta_t_tensor = res["delta_t_tensor"].to(s.device, s.delta_t_tensor.dtype).reshape(s.bs)
    return res


def precon_file_name(Lx, Ltau, dtau=0.1, t=1):
    """``precon_ckpt_L_{Lx}_Ltau_{Ltau}_dtau_{dtau}_t_{t}.pt`` (hmc_sampler_batch.py:463-465)."""
    return f"precon_ckpt_L_{Lx}_Ltau_{Ltau}_dtau_{dtau}_t_{t}.pt"


def save_precon(d, path):
    """Write a P dict ``{'indices', 'values', 'size'}`` (precon_manual.py:177-184)."""
    d = {"indices": d["indices"].detach().cpu(), "values": d["values"].detach().cpu().to(torch.complex64),
         "size": torch.Size(d["size"])}
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    torch.save(d, path)
    return path


def load_precon(path):
    """Read a P dict; checks the keys, index shape and value count."""
    d = _load(path)
    if not isinstance(d, dict) or not {"indices", "values", "size"} <= set(d):
        raise ValueError(f"{path}: not a preconditioner dict with 'indices', 'values' and 'size'")
    idx, val = d["indices"], d["values"]
    if idx.dim() != 2 or idx.shape[0] != 2 or val.dim() != 1 or val.shape[0] != idx.shape[1]:
        raise ValueError(f"{path}: indices {tuple(idx.shape)} and values {tuple(val.shape)} do not form a COO matrix")
    return {"indices": idx.to(torch.int64), "values": val, "size": torch.Size(d["size"])}


def precon_table_from_file(path, Lx, Ltau):
    """The band table of a cached P (float64 [Ltau, 13]) and its complex64 CSR.

    Raises QfError if the file's matrix is not P_tau (x) I_Vs with a 13-point
    cyclic tau band, the structure the kernels apply."""
    d = load_precon(path)
    n = Lx * Lx * Ltau
    if tuple(d["size"]) != (n, n):
        raise ValueError(f"{path}: size {tuple(d['size'])}, lattice needs ({n}, {n})")
    csr = precon.coo_to_csr(d, device="cpu")
    return precon.table_from_csr(csr, Lx * Lx, Ltau), csr
