"""The reference's tau-banded preconditioner P = P_tau (x) I_Vs.

``get_precon_man`` mirrors preconditioners_orig/precon_manual.py:15-184: it
returns the same COO dict ``{"indices", "values", "size"}`` (complex64) that the
reference saves and that ``HmcSampler.reset_precon`` (hmc_sampler_batch.py:
461-547) turns into ``precon_csr``.  On the GPU the operator is applied from its
band table (``table[tau, o + 6]`` multiplies ``r_{(tau+o) mod Ltau}``), which
the HIP library builds from the same float32 constants (qf_precon_table) or
extracts from a caller's CSR after verifying its structure
(qf_precon_table_from_csr).
"""
import numpy as np
import torch

from . import kernels


def band_table(Ltau, dtype=torch.float64, device="cpu"):
    """[Ltau, 13] band table of the manual preconditioner (needs Ltau >= 26)."""
    return kernels.precon_table(Ltau, dtype=dtype, device=device)


def table_to_coo(table, Vs):
    """Expand a band table to the reference's COO dict over N = Ltau*Vs sites."""
    tab = table.to("cpu", torch.float64).numpy()
    Lt = tab.shape[0]
    rows, cols, vals = [], [], []
    site = np.arange(Vs)
    for tau in range(Lt):
        for o in range(-6, 7):
            v = tab[tau, o + 6]
            if v == 0.0:
                continue
            tp = (tau + o) % Lt
            rows.append(tau * Vs + site)
            cols.append(tp * Vs + site)
            vals.append(np.full(Vs, v))
    idx = torch.from_numpy(np.stack([np.concatenate(rows), np.concatenate(cols)]))
    val = torch.from_numpy(np.concatenate(vals)).to(torch.complex64)
    coo = torch.sparse_coo_tensor(idx, val, (Lt * Vs, Lt * Vs)).coalesce()
    return {"indices": coo.indices(), "values": coo.values(), "size": coo.size()}


def get_precon_man(Lx, Ltau, save_path=None):
    """precon_manual.get_precon_man: the manual preconditioner as a COO dict (optionally saved with torch.save)."""
    d = table_to_coo(band_table(Ltau), Lx * Lx)
    if save_path is not None:
        torch.save(d, save_path)
    return d


def coo_to_csr(d, device="cuda"):
    """reset_precon's conversion (hmc_sampler_batch.py:532-547): COO dict -> complex64 CSR."""
    coo = torch.sparse_coo_tensor(d["indices"], d["values"], size=d["size"], dtype=torch.complex64).coalesce()
    return coo.to_sparse_csr().to(device)


def table_from_csr(P_csr, Vs, Ltau):
    """Verified band table of a caller's CSR preconditioner (raises QfError if it is not P_tau (x) I)."""
    return kernels.precon_table_from_csr(P_csr, Vs, Ltau)
