"""Drop-in for the reference's CUDA extension ``qed_fermion._C``.

The reference imports ``from qed_fermion import _C`` (hmc_sampler_batch.py:98-99,
stochastic_estimator.py:17-18) and calls four functions on torch tensors; the
extension itself is not in the reference tree, so its semantics are fixed by
those call sites.  Same names, arguments and return values here; the work runs
in the HIP kernels of libqf_hip.so.  Like the reference, inputs are not mutated
and fresh output tensors are returned.  ``bx, by`` (the reference's
BLOCK_SIZE=(4, 8)) are accepted and ignored: the MI355X tiling is chosen by the
library (qf_set_tuning / autotune).

Dtypes: complex64 vectors with float32 boson (the reference's), or complex128
with float64 (the FP64 path).
"""
import torch

from . import kernels
from ._lib import QF_ERR_INVALID, QfError


def _ltau(x, Lx):
    N = x.shape[-1]
    if N % (Lx * Lx):
        raise QfError(QF_ERR_INVALID, f"vector length {N} is not a multiple of Lx^2={Lx * Lx}")
    return N // (Lx * Lx)


def _boson(boson, x):
    want = torch.float32 if x.dtype == torch.complex64 else torch.float64
    if boson.dtype != want:
        raise QfError(QF_ERR_INVALID, f"boson dtype {boson.dtype} does not match vector dtype {x.dtype}")
    return boson.contiguous()


def mhm_vec(boson, x, Lx, dtau, bx=4, by=8):
    """y = M^dag M x, batched: boson [bs, Ltau*Ly*Lx*2], x [bs, Ltau*Ly*Lx] (hmc_sampler_batch.py:772, 799)."""
    x = x.contiguous()
    return kernels.apply("MdagM", _boson(boson, x), x, Lx, _ltau(x, Lx), dtau)


def mh_vec(boson, x, Lx, dtau, bx=4, by=8):
    """y = M^dag x, batched (hmc_sampler_batch.py:2406; stochastic_estimator.py:388, 430)."""
    x = x.contiguous()
    return kernels.apply("Mdag", _boson(boson, x), x, Lx, _ltau(x, Lx), dtau)


_table_cache = {}


def _table_of(P_csr, Vs, Ltau, dtype, device):
    key = (id(P_csr), P_csr.values().data_ptr(), Vs, Ltau, dtype, str(device))
    tab = _table_cache.get(key)
    if tab is None:
        tab = kernels.precon_table_from_csr(P_csr, Vs, Ltau).to(device=device, dtype=dtype)
        _table_cache.clear()
        _table_cache[key] = tab
    return tab


def precon_vec(r, precon_csr, Lx):
    """z = P r with the reference's CSR preconditioner (hmc_sampler_batch.py:773, 838).

    The CSR is verified once to be P_tau (x) I_Vs with a 13-point tau band
    (otherwise QfError QF_ERR_PRECON) and applied from its band table."""
    r = r.contiguous()
    Vs = Lx * Lx
    Ltau = _ltau(r, Lx)
    real = torch.float32 if r.dtype == torch.complex64 else torch.float64
    tab = _table_of(precon_csr, Vs, Ltau, real, r.device)
    return kernels.apply_precon(tab, r, Vs, Ltau)


def b_vec_per_tau(boson_tau, v, Lx, dtau, interm, bx=4, by=8):
    """B_tau applied to one slice (hmc_sampler_batch.py:2181-2189).

    boson_tau [Ly*Lx*2], v [Vs].  interm=False: returns B_tau v ([Vs]).
    interm=True: returns [6, Vs] partial products in the order force_f_fast
    consumes them (its group-g gathers pair entry k with the sparse path's
    xi_c_rgt_k, 2260-2325): out[k] = Q_{5-k} v for k = 0..4 and out[5] = Q_6 v,
    where Q_m applies the first m factors of D4 D3 D2 D1 D2 D3 D4."""
    v = v.contiguous().view(1, -1)
    b = _boson(boson_tau, v).view(1, -1)
    if not interm:
        return kernels.apply("B", b, v, Lx, 1, dtau).view(-1)
    outs = [kernels.apply_partial_B(b, v, Lx, 1, dtau, m).view(-1) for m in (5, 4, 3, 2, 1, 6)]
    return torch.stack(outs)
