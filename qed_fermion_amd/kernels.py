"""Torch-tensor front end of the HIP kernels in libqf_hip.so.

Tensors must be contiguous CUDA (HIP) tensors in the reference's permuted
layouts (hmc_sampler_batch.py:2166): ``phi`` ``[nchain, Ltau*L*L*2]`` real and
vectors ``[nrhs, Ltau*L*L]`` complex.  complex64/float32 selects the FP32
kernels, complex128/float64 the FP64 ones.  Every launch goes on torch's
current stream, so these calls can be captured with ``torch.cuda.graph``.
Outputs and workspaces are allocated with torch (never inside the library).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import QfError, call, ptr, stream_handle

_PREC = {torch.complex64: _lib.QF_PREC_F32, torch.complex128: _lib.QF_PREC_F64,
         torch.float32: _lib.QF_PREC_F32, torch.float64: _lib.QF_PREC_F64}
_REAL = {torch.complex64: torch.float32, torch.complex128: torch.float64}


def _prec(t):
    try:
        return _PREC[t.dtype]
    except KeyError:
        raise QfError(_lib.QF_ERR_INVALID, f"unsupported dtype {t.dtype}") from None


def _check_dev(*ts):
    for t in ts:
        if t is None:
            continue
        if not t.is_cuda:
            raise QfError(_lib.QF_ERR_INVALID, "tensor is not on the GPU; the HIP path has no CPU fallback")
        if not t.is_contiguous():
            raise QfError(_lib.QF_ERR_INVALID, "tensor must be contiguous")


def _buf(t, name, dtype, numel, device=None):
    """Check a caller-supplied buffer before its pointer is taken: a GPU tensor (on ``device``
    when given) of exactly ``numel`` elements of ``dtype``, contiguous -- so that the kernel
    neither reads or writes past it nor writes into a temporary copy."""
    if t is None:
        raise QfError(_lib.QF_ERR_INVALID, f"{name}: a tensor is required")
    if t.dtype != dtype:
        raise QfError(_lib.QF_ERR_INVALID, f"{name} has dtype {t.dtype}, expected {dtype}")
    if t.numel() != numel:
        raise QfError(_lib.QF_ERR_INVALID, f"{name} has {t.numel()} elements, expected {numel}")
    if not t.is_contiguous():
        raise QfError(_lib.QF_ERR_INVALID, f"{name} must be contiguous")
    if not t.is_cuda:
        raise QfError(_lib.QF_ERR_INVALID, f"{name} is not on the GPU; the HIP path has no CPU fallback")
    if device is not None and t.device != device:
        raise QfError(_lib.QF_ERR_INVALID, f"{name} is on {t.device}, expected {device}")
    return t


def _vecs(x, N, name="x"):
    """[nrhs, N] view of a contiguous GPU vector batch of complex64/complex128."""
    if x.dtype not in _REAL:
        raise QfError(_lib.QF_ERR_INVALID, f"{name} has dtype {x.dtype}, expected complex64 or complex128")
    if x.numel() == 0 or x.numel() % N:
        raise QfError(_lib.QF_ERR_INVALID, f"{name} has {x.numel()} elements, not a positive multiple of N={N}")
    _check_dev(x)
    return x.view(-1, N)


def _work(work, nbytes, device):
    if work.dtype != torch.uint8 or not work.is_cuda or work.device != device or not work.is_contiguous():
        raise QfError(_lib.QF_ERR_INVALID, "workspace must be a contiguous uint8 tensor on the solve's device")
    if work.numel() < nbytes:
        raise QfError(_lib.QF_ERR_INVALID, f"workspace of {work.numel()} bytes, need {nbytes}")
    return work


def _phi_stride(phi, nrhs, N):
    """Chain stride of phi: 0 when one field is shared by all right-hand sides."""
    nphi = phi.numel() // (2 * N)
    if phi.numel() != nphi * 2 * N:
        raise QfError(_lib.QF_ERR_INVALID, f"phi has {phi.numel()} reals, not a multiple of 2*N={2 * N}")
    if nphi == nrhs:
        return 2 * N
    if nphi == 1:
        return 0
    raise QfError(_lib.QF_ERR_INVALID, f"phi holds {nphi} fields for {nrhs} right-hand sides")


def apply(op, phi, x, L, Ltau, dtau=0.1, t=1.0, out=None):
    """y = op(phi) x; op in {'M', 'Mdag', 'MdagM', 'B'}."""
    code = {"M": _lib.QF_OP_M, "Mdag": _lib.QF_OP_MDAG, "MdagM": _lib.QF_OP_MDAGM, "B": _lib.QF_OP_B}[op]
    N = L * L * Ltau
    x2 = _vecs(x, N)
    _check_dev(phi)
    if phi.dtype != _REAL[x.dtype]:
        raise QfError(_lib.QF_ERR_INVALID, f"phi dtype {phi.dtype} does not match vector dtype {x.dtype}")
    nrhs = x2.shape[0]
    y = torch.empty_like(x2) if out is None else _buf(out, "out", x.dtype, x2.numel(), x.device).view(-1, N)
    call("qf_apply", code, _prec(x), ptr(phi), _phi_stride(phi, nrhs, N), ptr(x2), ptr(y), nrhs, L, Ltau, float(dtau),
         float(t), stream_handle())
    return y.view(x.shape)


def hop_table(phi, L, Ltau, dtau=0.1, t=1.0, out=None):
    """Per-link factor coefficients tanh(a_g) e^{i phi} of each field in phi [nchain, 2N]."""
    N = L * L * Ltau
    _check_dev(phi)
    if phi.dtype not in (torch.float32, torch.float64) or phi.numel() == 0 or phi.numel() % (2 * N):
        raise QfError(_lib.QF_ERR_INVALID, f"phi must be float32/float64 with a multiple of 2N={2 * N} elements")
    nchain = phi.numel() // (2 * N)
    cd = torch.complex64 if phi.dtype == torch.float32 else torch.complex128
    hop = torch.empty(nchain, 2 * N, dtype=cd, device=phi.device) if out is None else _buf(out, "out", cd, nchain * 2 * N,
                                                                                            phi.device)
    call("qf_hop_table", _prec(phi), ptr(phi), 2 * N, nchain, L, Ltau, float(dtau), float(t), ptr(hop), stream_handle())
    return hop


def apply_hop(op, hop, x, L, Ltau, dtau=0.1, t=1.0, out=None):
    """y = op x with the operator given by its hop table (hop [nchain, 2N] complex; nchain 1 = shared)."""
    code = {"M": _lib.QF_OP_M, "Mdag": _lib.QF_OP_MDAG, "MdagM": _lib.QF_OP_MDAGM, "B": _lib.QF_OP_B}[op]
    N = L * L * Ltau
    x2 = _vecs(x, N)
    _check_dev(hop)
    if hop.dtype != x.dtype:
        raise QfError(_lib.QF_ERR_INVALID, f"hop table dtype {hop.dtype} does not match vector dtype {x.dtype}")
    nrhs = x2.shape[0]
    nchain = hop.numel() // (2 * N)
    hop_cs = 0 if nchain == 1 else 2 * N
    if nchain not in (1, nrhs) or hop.numel() != nchain * 2 * N:
        raise QfError(_lib.QF_ERR_INVALID, f"hop table of {hop.numel()} entries for {nrhs} right-hand sides")
    y = torch.empty_like(x2) if out is None else _buf(out, "out", x.dtype, x2.numel(), x.device).view(-1, N)
    call("qf_apply_hop", code, _prec(x), ptr(hop), hop_cs, ptr(x2), ptr(y), nrhs, L, Ltau, float(dtau), float(t),
         stream_handle())
    return y.view(x.shape)


def apply_partial_B(phi, x, L, Ltau, dtau, nfac, t=1.0):
    """y_tau = Q_nfac x_tau: the first nfac factors of B_tau = D4 D3 D2 D1 D2 D3 D4 applied per slice."""
    if not 0 <= nfac <= 7:
        raise QfError(_lib.QF_ERR_INVALID, f"nfac={nfac}")
    N = L * L * Ltau
    x2 = _vecs(x, N)
    _check_dev(phi)
    if phi.dtype != _REAL[x.dtype]:
        raise QfError(_lib.QF_ERR_INVALID, f"phi dtype {phi.dtype} does not match vector dtype {x.dtype}")
    y = torch.empty_like(x2)
    call("qf_apply", 16 + int(nfac), _prec(x), ptr(phi), _phi_stride(phi, x2.shape[0], N), ptr(x2), ptr(y), x2.shape[0], L,
         Ltau, float(dtau), float(t), stream_handle())
    return y.view(x.shape)


def precon_table(Ltau, dtype=torch.float64, device="cuda"):
    """The manual preconditioner (precon_manual.py:15-175) as a [Ltau, 13] band table."""
    tab = np.zeros((Ltau, 13), dtype=np.float64)
    call("qf_precon_table", Ltau, tab.ctypes.data_as(ctypes.c_void_p))
    return torch.from_numpy(tab).to(device=device, dtype=dtype)


def precon_table_from_csr(P_csr, Vs, Ltau):
    """Verify that a torch CSR matrix is P_tau (x) I_Vs with a 13-band tau stencil; return its table (float64, CPU)."""
    P = P_csr.to("cpu")
    crow = P.crow_indices().to(torch.int64).numpy()
    col = P.col_indices().to(torch.int64).numpy()
    vals = P.values().detach().clone()
    if vals.is_complex():
        ri = torch.view_as_real(vals).to(torch.float64).numpy()
        vr, vi = np.ascontiguousarray(ri[:, 0]), np.ascontiguousarray(ri[:, 1])
    else:
        vr, vi = np.ascontiguousarray(vals.to(torch.float64).numpy()), None
    crow = np.ascontiguousarray(crow)
    col = np.ascontiguousarray(col)
    tab = np.zeros((Ltau, 13), dtype=np.float64)
    c_p = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
    call("qf_precon_table_from_csr", c_p(crow), c_p(col), c_p(vr), c_p(vi), P.shape[0], Vs, Ltau, c_p(tab))
    return torch.from_numpy(tab)


def apply_precon(ptab, r, Vs, Ltau, out=None):
    N = Vs * Ltau
    r2 = _vecs(r, N, "r")
    _check_dev(ptab)
    if ptab.numel() != 13 * Ltau:
        raise QfError(_lib.QF_ERR_INVALID, f"preconditioner table has {ptab.numel()} entries, expected 13*Ltau={13 * Ltau}")
    tab = ptab.to(_REAL[r.dtype]).contiguous()
    z = torch.empty_like(r2) if out is None else _buf(out, "out", r.dtype, r2.numel(), r.device).view(-1, N)
    call("qf_precon_vec", _prec(r), ptr(tab), ptr(r2), ptr(z), r2.shape[0], Vs, Ltau, stream_handle())
    return z.view(r.shape)


_work_cache = {}


def _workspace(key, nbytes, device):
    buf = _work_cache.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
        _work_cache[key] = buf
    return buf


def pcg_workspace_bytes(prec, mixed, nrhs, L, Ltau):
    return _lib.lib().qf_pcg_workspace_bytes(prec, int(bool(mixed)), nrhs, L, Ltau)


def _status_buf(status, nrhs, dev):
    return None if status is None else _buf(status, "status", torch.int32, nrhs, dev)


def check_status(status, what="PCG solve", allow_max_iter=True):
    """Raise QfError if a chain's solve failed (a QF_SOLVE_* code of NONFINITE or BREAKDOWN, or
    MAX_ITER when not allowed).  Reads the device status: a host synchronisation, so not for
    code under graph capture."""
    codes = status.tolist()
    bad = [(i, c) for i, c in enumerate(codes)
           if c in (_lib.QF_SOLVE_NONFINITE, _lib.QF_SOLVE_BREAKDOWN, _lib.QF_SOLVE_ABORTED) or (c == _lib.QF_SOLVE_MAX_ITER and not allow_max_iter)]
    if bad:
        i, c = bad[0]
        raise QfError(_lib.QF_ERR_INVALID, f"{what}: chain {i}: {_lib.SOLVE_STATUS_NAMES.get(c, c)} "
                                           f"({len(bad)} of {len(codes)} chains failed)")
    return codes


def pcg_solve(phi, b, L, Ltau, dtau=0.1, t=1.0, ptab=None, rtol=1e-9, max_iter=400, x=None, work=None, err=None, iters=None,
              status=None):
    """Batched PCG (preconditioned_cg_fast_test recurrence).  Returns (x, iters, err) as device tensors;
    ``status`` (int32 [nrhs], optional) receives each chain's QF_SOLVE_* code."""
    N = L * L * Ltau
    b2 = _vecs(b, N, "b")
    _check_dev(phi, ptab)
    if phi.dtype != _REAL[b.dtype]:
        raise QfError(_lib.QF_ERR_INVALID, f"phi dtype {phi.dtype} does not match vector dtype {b.dtype}")
    nrhs = b2.shape[0]
    prec = _prec(b)
    dev = b.device
    x = torch.empty_like(b2) if x is None else _buf(x, "x", b.dtype, b2.numel(), dev).view(-1, N)
    nbytes = pcg_workspace_bytes(prec, 0, nrhs, L, Ltau)
    if nbytes == 0:
        raise QfError(_lib.QF_ERR_INVALID, _lib.lib().qf_last_error().decode())
    work = _workspace(("pcg", prec, nrhs, L, Ltau, dev), nbytes, dev) if work is None else _work(work, nbytes, dev)
    err = torch.empty(nrhs, dtype=torch.float64, device=dev) if err is None else _buf(err, "err", torch.float64, nrhs, dev)
    iters = torch.empty(nrhs, dtype=torch.float64, device=dev) if iters is None else _buf(iters, "iters", torch.float64, nrhs, dev)
    if ptab is not None and ptab.numel() != 13 * Ltau:
        raise QfError(_lib.QF_ERR_INVALID, f"preconditioner table has {ptab.numel()} entries, expected 13*Ltau={13 * Ltau}")
    tab = ptab.to(_REAL[b.dtype]).contiguous() if ptab is not None else None
    st = _status_buf(status, nrhs, dev)
    call("qf_pcg_solve", prec, ptr(phi), _phi_stride(phi, nrhs, N), ptr(b2), ptr(x), nrhs, L, Ltau, float(dtau), float(t),
         ptr(tab), float(rtol), int(max_iter), ptr(work), work.numel(), ptr(err), ptr(iters), ptr(st), stream_handle())
    return x.view(b.shape), iters, err


def pcg_persistent_workspace_bytes(prec, nrhs, L, Ltau):
    return _lib.lib().qf_pcg_persistent_workspace_bytes(prec, nrhs, L, Ltau)


def pcg_solve_persistent(phi, b, L, Ltau, dtau=0.1, t=1.0, ptab=None, rtol=1e-9, max_iter=400, x=None, err=None, iters=None,
                         status=None):
    """pcg_solve as ONE persistent launch (qf_pcg_solve_persistent): the same recurrence and
    outputs for any number of right-hand sides in a single dispatch, leaving as soon as no chain
    is active.  ptab is required.  Returns (x, iters, err) as device tensors; ``status`` (int32
    [nrhs], optional) receives each chain's QF_SOLVE_* code (QF_SOLVE_ABORTED if a grid barrier
    timed out).  Not for graph capture."""
    N = L * L * Ltau
    b2 = _vecs(b, N, "b")
    if ptab is None:
        raise QfError(_lib.QF_ERR_INVALID, "the persistent solve needs the preconditioner table")
    _check_dev(phi, ptab)
    if phi.dtype != _REAL[b.dtype]:
        raise QfError(_lib.QF_ERR_INVALID, f"phi dtype {phi.dtype} does not match vector dtype {b.dtype}")
    if ptab.numel() != 13 * Ltau:
        raise QfError(_lib.QF_ERR_INVALID, f"preconditioner table has {ptab.numel()} entries, expected 13*Ltau={13 * Ltau}")
    nrhs = b2.shape[0]
    prec = _prec(b)
    dev = b.device
    x = torch.empty_like(b2) if x is None else _buf(x, "x", b.dtype, b2.numel(), dev).view(-1, N)
    nbytes = pcg_persistent_workspace_bytes(prec, nrhs, L, Ltau)
    if nbytes == 0:
        raise QfError(_lib.QF_ERR_UNSUPPORTED, _lib.lib().qf_last_error().decode())
    # the barrier counters at the end of the workspace start at zero and every solve leaves them zero
    key = ("persistent", prec, nrhs, L, Ltau, dev)
    work = _work_cache.get(key)
    if work is None or work.numel() != nbytes:
        work = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
        _work_cache[key] = work
    err = torch.empty(nrhs, dtype=torch.float64, device=dev) if err is None else _buf(err, "err", torch.float64, nrhs, dev)
    iters = torch.empty(nrhs, dtype=torch.float64, device=dev) if iters is None else _buf(iters, "iters", torch.float64, nrhs, dev)
    tab = ptab.to(_REAL[b.dtype]).contiguous()
    st = _status_buf(status, nrhs, dev)
    call("qf_pcg_solve_persistent", prec, ptr(phi), _phi_stride(phi, nrhs, N), ptr(b2), ptr(x), nrhs, L, Ltau, float(dtau),
         float(t), ptr(tab), float(rtol), int(max_iter), ptr(work), work.numel(), ptr(err), ptr(iters), ptr(st), stream_handle())
    return x.view(b.shape), iters, err


def pcg_mixed_solve(phi64, b, L, Ltau, dtau=0.1, t=1.0, ptab=None, rtol=1e-10, max_iter=400, ru_period=50, phi32=None,
                    x=None, work=None, err=None, iters=None, status=None):
    """Mixed-precision PCG: complex64 iterations with FP64 reliable updates every ``ru_period``
    iterations.  b complex128, phi64 float64.  Returns (x complex128, iters, err_true);
    ``status`` (int32 [nrhs], optional) receives each chain's QF_SOLVE_* code."""
    N = L * L * Ltau
    b2 = _vecs(b, N, "b")
    if b2.dtype != torch.complex128 or phi64.dtype != torch.float64:
        raise QfError(_lib.QF_ERR_INVALID, "mixed solve takes complex128 b and float64 phi")
    _check_dev(phi64, ptab)
    nrhs = b2.shape[0]
    dev = b.device
    _phi_stride(phi64, nrhs, N)
    phi32 = phi64.to(torch.float32) if phi32 is None else _buf(phi32, "phi32", torch.float32, phi64.numel(), dev)
    x = torch.empty_like(b2) if x is None else _buf(x, "x", torch.complex128, b2.numel(), dev).view(-1, N)
    nbytes = pcg_workspace_bytes(_lib.QF_PREC_F32, 1, nrhs, L, Ltau)
    if nbytes == 0:
        raise QfError(_lib.QF_ERR_INVALID, _lib.lib().qf_last_error().decode())
    work = _workspace(("mixed", nrhs, L, Ltau, dev), nbytes, dev) if work is None else _work(work, nbytes, dev)
    err = torch.empty(nrhs, dtype=torch.float64, device=dev) if err is None else _buf(err, "err", torch.float64, nrhs, dev)
    iters = torch.empty(nrhs, dtype=torch.float64, device=dev) if iters is None else _buf(iters, "iters", torch.float64, nrhs, dev)
    if ptab is not None and ptab.numel() != 13 * Ltau:
        raise QfError(_lib.QF_ERR_INVALID, f"preconditioner table has {ptab.numel()} entries, expected 13*Ltau={13 * Ltau}")
    tab = ptab.to(torch.float32).contiguous() if ptab is not None else None
    st = _status_buf(status, nrhs, dev)
    call("qf_pcg_mixed_solve", ptr(phi64), ptr(phi32), _phi_stride(phi64, nrhs, N), ptr(b2), ptr(x), nrhs, L, Ltau,
         float(dtau), float(t), ptr(tab), float(rtol), int(max_iter), int(ru_period), ptr(work), work.numel(), ptr(err),
         ptr(iters), ptr(st), stream_handle())
    return x.view(b.shape), iters, err


def fermion_force(phi, xi, L, Ltau, dtau=0.1, t=1.0, out=None):
    """Ft = -dS_f/dphi in the phi layout [nrhs, Ltau*L*L*2]."""
    N = L * L * Ltau
    xi2 = _vecs(xi, N, "xi")
    _check_dev(phi)
    if phi.dtype != _REAL[xi.dtype]:
        raise QfError(_lib.QF_ERR_INVALID, f"phi dtype {phi.dtype} does not match vector dtype {xi.dtype}")
    nrhs = xi2.shape[0]
    rd = _REAL[xi.dtype]
    F = torch.empty(nrhs, 2 * N, dtype=rd, device=xi.device) if out is None else _buf(out, "out", rd, nrhs * 2 * N, xi.device)
    call("qf_fermion_force", _prec(xi), ptr(phi), _phi_stride(phi, nrhs, N), ptr(xi2), ptr(F), nrhs, L, Ltau, float(dtau),
         float(t), stream_handle())
    return F


def _fields64(x, L, Ltau, name="x"):
    """Number of float64 fields [nrhs, 2*Ltau*L*L] in x (checked)."""
    n2 = 2 * L * L * Ltau
    if x.dtype != torch.float64 or x.numel() == 0 or x.numel() % n2:
        raise QfError(_lib.QF_ERR_INVALID, f"{name} must be float64 with a positive multiple of 2N={n2} elements")
    _check_dev(x)
    return x.numel() // n2


def boson_force(x, L, Ltau, cJ, cK, compact=False, Fp=None, Ft=None):
    nrhs = _fields64(x, L, Ltau)
    Fp = torch.empty_like(x) if Fp is None else _buf(Fp, "Fp", torch.float64, x.numel(), x.device)
    Ft = torch.empty_like(x) if Ft is None else _buf(Ft, "Ft", torch.float64, x.numel(), x.device)
    call("qf_boson_force", ptr(x), ptr(Fp), ptr(Ft), nrhs, L, Ltau, float(cJ), float(cK), int(bool(compact)), stream_handle())
    return Fp, Ft


def leapfrog_boson(x, p, Fp, Ft, dt, L, Ltau, cJ, cK, compact=False, m=1.0, nsub=5, tau_mask=None):
    """leapfrog_cmp in place on (x, p, Fp, Ft); dt float64 [nchain]."""
    nrhs = _fields64(x, L, Ltau)
    for t, name in ((p, "p"), (Fp, "Fp"), (Ft, "Ft")):
        _buf(t, name, torch.float64, x.numel(), x.device)
    _buf(dt, "dt", torch.float64, nrhs, x.device)
    if tau_mask is not None:
        _buf(tau_mask, "tau_mask", torch.float64, Ltau, x.device)
    call("qf_leapfrog_boson", ptr(x), ptr(p), ptr(Fp), ptr(Ft), ptr(dt), ptr(tau_mask), float(m), int(nsub), nrhs, L, Ltau,
         float(cJ), float(cK), int(bool(compact)), stream_handle())


def momentum_kick(p, F, dt, scale, L, Ltau, tau_mask=None):
    nrhs = _fields64(p, L, Ltau, "p")
    _buf(F, "F", torch.float64, p.numel(), p.device)
    _buf(dt, "dt", torch.float64, nrhs, p.device)
    if tau_mask is not None:
        _buf(tau_mask, "tau_mask", torch.float64, Ltau, p.device)
    call("qf_momentum_kick", ptr(p), ptr(F), ptr(dt), float(scale), ptr(tau_mask), nrhs, L, Ltau, stream_handle())


def hamiltonian(x, L, Ltau, cJ, cK, compact=False, m=1.0, p=None, psi=None, xi=None, H=None, parts=None):
    """Per-chain H = S_tau + S_plaq + p^2/2m + Re(psi^dag xi) in FP64; returns (H, parts[nrhs, 4])."""
    nrhs = _fields64(x, L, Ltau)
    dev = x.device
    if p is not None:
        _buf(p, "p", torch.float64, x.numel(), dev)
    if (psi is None) != (xi is None):
        raise QfError(_lib.QF_ERR_INVALID, "psi and xi must both be given or both be None")
    if psi is not None:
        if psi.dtype not in _REAL:
            raise QfError(_lib.QF_ERR_INVALID, f"psi has dtype {psi.dtype}, expected complex64 or complex128")
        _buf(psi, "psi", psi.dtype, nrhs * L * L * Ltau, dev)
        _buf(xi, "xi", psi.dtype, nrhs * L * L * Ltau, dev)
    H = torch.empty(nrhs, dtype=torch.float64, device=dev) if H is None else _buf(H, "H", torch.float64, nrhs, dev)
    parts = torch.empty(nrhs, 4, dtype=torch.float64, device=dev) if parts is None else _buf(parts, "parts", torch.float64,
                                                                                              4 * nrhs, dev)
    prec = _prec(psi) if psi is not None else _lib.QF_PREC_F64
    nbytes = _lib.lib().qf_hamiltonian_workspace_bytes(nrhs)
    work = _workspace(("ham", nrhs, x.device), nbytes, x.device)
    call("qf_hamiltonian", prec, ptr(x), ptr(p), ptr(psi), ptr(xi), ptr(H), ptr(parts), nrhs, L, Ltau, float(cJ), float(cK),
         int(bool(compact)), float(m), ptr(work), work.numel(), stream_handle())
    return H, parts


def set_tuning(prec, by=0, t=0, tb=0):
    call("qf_set_tuning", int(prec), int(by), int(t), int(tb))


def set_tuning_shape(prec, L, Ltau, nrhs=0, by=0, t=0, tb=0):
    """Tiling for one lattice shape (nrhs=0: any batch size); see autotune.py."""
    call("qf_set_tuning_shape", int(prec), int(L), int(Ltau), int(nrhs), int(by), int(t), int(tb))


def clear_tuning():
    call("qf_clear_tuning")


def get_tiling(prec, nrhs, L, Ltau):
    out = (ctypes.c_int * 4)()
    call("qf_get_tiling", int(prec), nrhs, L, Ltau, out)
    ns = ctypes.c_int()
    call("qf_get_rhs_per_group", int(prec), nrhs, L, Ltau, ctypes.byref(ns))
    return {"by": out[0], "t": out[1], "tb": out[2], "nA": out[3], "ns": ns.value}


def set_tuning_ns(prec, L, Ltau, nrhs=0, ns=0):
    """Right-hand sides per work-group of the plain M^dag M from phi for one shape (0 or 1:
    one); several fill whole wavefronts where one slice does not (see autotune.py)."""
    call("qf_set_tuning_ns", int(prec), int(L), int(Ltau), int(nrhs), int(ns))


# ------------------------------------------------------------------ ragged batch
class RaggedPlan:
    """Plan of a ragged batch: independent lattices of different L (one chain
    each) solved in the same launches.  Holds the host plan and its device copy
    (the C ABI takes both)."""

    def __init__(self, Ls, Ltaus, device="cuda"):
        import numpy as np
        self.Ls = [int(v) for v in Ls]
        self.Ltaus = [int(v) for v in Ltaus]
        if len(self.Ls) != len(self.Ltaus) or not self.Ls:
            raise QfError(_lib.QF_ERR_INVALID, "L and Ltau lists must be non-empty and of equal length")
        n = len(self.Ls)
        nbytes = _lib.lib().qf_ragged_plan_bytes(n)
        self.host = np.zeros(nbytes, dtype=np.uint8)
        Li = np.asarray(self.Ls, dtype=np.int32)
        Lti = np.asarray(self.Ltaus, dtype=np.int32)
        call("qf_ragged_plan", n, Li.ctypes.data, Lti.ctypes.data, self.host.ctypes.data, nbytes)
        self.dev = torch.from_numpy(self.host.copy()).to(device)
        self.sizes = [L * L * Lt for L, Lt in zip(self.Ls, self.Ltaus)]
        self.offsets = [0]
        for v in self.sizes:
            self.offsets.append(self.offsets[-1] + v)
        self.nsites = self.offsets[-1]

    def split(self, v):
        """Per-lattice views of a concatenated vector (complex: one value per site)."""
        return [v[self.offsets[k]:self.offsets[k + 1]] for k in range(len(self.Ls))]

    def precon_table(self, dtype=torch.float64, device="cuda"):
        """The manual preconditioner's band tables of every lattice, concatenated."""
        return torch.cat([precon_table(Lt, dtype=dtype, device=device).reshape(-1) for Lt in self.Ltaus])


def apply_ragged(plan, phi, x, dtau=0.1, t=1.0, out=None):
    """y = M^dag M x for every lattice of a ragged batch in one launch.
    phi: concatenated [2 * sum N_k] real; x: [sum N_k] complex."""
    _check_dev(phi, x)
    if x.numel() != plan.nsites or phi.numel() != 2 * plan.nsites:
        raise QfError(_lib.QF_ERR_INVALID, "ragged vector / field sizes do not match the plan")
    if x.dtype not in _REAL or phi.dtype != _REAL[x.dtype]:
        raise QfError(_lib.QF_ERR_INVALID, f"ragged dtypes: x {x.dtype}, phi {phi.dtype}")
    y = torch.empty_like(x) if out is None else _buf(out, "out", x.dtype, x.numel(), x.device)
    call("qf_apply_ragged", _lib.QF_OP_MDAGM, _prec(x), plan.host.ctypes.data, ptr(plan.dev), ptr(phi), ptr(x), ptr(y),
         float(dtau), float(t), stream_handle())
    return y


def pcg_solve_ragged(plan, phi, b, dtau=0.1, t=1.0, ptab=None, rtol=1e-9, max_iter=400, x=None, status=None):
    """Batched PCG over a ragged batch.  Returns (x, iters, err) with one
    iters / err entry per lattice; ``status`` (int32 [lattices], optional) gets the QF_SOLVE_* codes."""
    _check_dev(phi, b, ptab)
    if b.numel() != plan.nsites or phi.numel() != 2 * plan.nsites:
        raise QfError(_lib.QF_ERR_INVALID, "ragged vector / field sizes do not match the plan")
    if b.dtype not in _REAL or phi.dtype != _REAL[b.dtype]:
        raise QfError(_lib.QF_ERR_INVALID, f"ragged dtypes: b {b.dtype}, phi {phi.dtype}")
    if ptab is not None and ptab.numel() != 13 * sum(plan.Ltaus):
        raise QfError(_lib.QF_ERR_INVALID, "preconditioner tables do not match the plan")
    prec = _prec(b)
    nbytes = _lib.lib().qf_pcg_ragged_workspace_bytes(prec, plan.host.ctypes.data)
    work = _workspace(("ragged", prec, tuple(plan.Ls), tuple(plan.Ltaus), b.device), nbytes, b.device)
    x = torch.empty_like(b) if x is None else _buf(x, "x", b.dtype, b.numel(), b.device)
    n = len(plan.Ls)
    err = torch.empty(n, dtype=torch.float64, device=b.device)
    iters = torch.empty(n, dtype=torch.float64, device=b.device)
    tab = ptab.to(_REAL[b.dtype]).contiguous() if ptab is not None else None
    st = _status_buf(status, n, b.device)
    call("qf_pcg_solve_ragged", prec, plan.host.ctypes.data, ptr(plan.dev), ptr(phi), ptr(b), ptr(x), float(dtau), float(t),
         ptr(tab), float(rtol), int(max_iter), ptr(work), work.numel(), ptr(err), ptr(iters), ptr(st), stream_handle())
    return x, iters, err


def pcg_solve_ragged_persistent(plan, phi, b, dtau=0.1, t=1.0, ptab=None, rtol=1e-9, max_iter=400, x=None, status=None):
    """pcg_solve_ragged as ONE persistent launch (qf_pcg_solve_ragged_persistent), leaving as
    soon as every lattice has stopped.  ptab (the plan's concatenated tables) is required.
    Returns (x, iters, err) with one iters / err entry per lattice; ``status`` (int32
    [lattices], optional) gets the QF_SOLVE_* codes.  Not for graph capture."""
    _check_dev(phi, b, ptab)
    if ptab is None:
        raise QfError(_lib.QF_ERR_INVALID, "the persistent solve needs the preconditioner tables")
    if b.numel() != plan.nsites or phi.numel() != 2 * plan.nsites:
        raise QfError(_lib.QF_ERR_INVALID, "ragged vector / field sizes do not match the plan")
    if b.dtype not in _REAL or phi.dtype != _REAL[b.dtype]:
        raise QfError(_lib.QF_ERR_INVALID, f"ragged dtypes: b {b.dtype}, phi {phi.dtype}")
    if ptab.numel() != 13 * sum(plan.Ltaus):
        raise QfError(_lib.QF_ERR_INVALID, "preconditioner tables do not match the plan")
    prec = _prec(b)
    nbytes = _lib.lib().qf_pcg_ragged_persistent_workspace_bytes(prec, plan.host.ctypes.data)
    key = ("ragged_persistent", prec, tuple(plan.Ls), tuple(plan.Ltaus), b.device)
    work = _work_cache.get(key)
    if work is None or work.numel() != nbytes:   # barrier counters start at zero; every solve leaves them zero
        work = torch.zeros(nbytes, dtype=torch.uint8, device=b.device)
        _work_cache[key] = work
    x = torch.empty_like(b) if x is None else _buf(x, "x", b.dtype, b.numel(), b.device)
    n = len(plan.Ls)
    err = torch.empty(n, dtype=torch.float64, device=b.device)
    iters = torch.empty(n, dtype=torch.float64, device=b.device)
    tab = ptab.to(_REAL[b.dtype]).contiguous()
    st = _status_buf(status, n, b.device)
    call("qf_pcg_solve_ragged_persistent", prec, plan.host.ctypes.data, ptr(plan.dev), ptr(phi), ptr(b), ptr(x), float(dtau),
         float(t), ptr(tab), float(rtol), int(max_iter), ptr(work), work.numel(), ptr(err), ptr(iters), ptr(st), stream_handle())
    return x, iters, err


def pcg_mixed_solve_ragged(plan, phi64, b, dtau=0.1, t=1.0, ptab=None, rtol=1e-10, max_iter=400, ru_period=50, phi32=None,
                           x=None, status=None):
    """Mixed-precision PCG over a ragged batch (complex64 iterations, FP64 reliable updates at most
    every ``ru_period`` iterations).  phi64 float64 [2 * sum N_k], b complex128 [sum N_k].  Returns
    (x complex128, iters, err_true) with one iters / err entry per lattice; ``status`` (int32
    [lattices], optional) gets the QF_SOLVE_* codes."""
    _check_dev(phi64, b, ptab)
    if b.numel() != plan.nsites or phi64.numel() != 2 * plan.nsites:
        raise QfError(_lib.QF_ERR_INVALID, "ragged vector / field sizes do not match the plan")
    if b.dtype != torch.complex128 or phi64.dtype != torch.float64:
        raise QfError(_lib.QF_ERR_INVALID, "the mixed ragged solve takes complex128 b and float64 phi")
    if ptab is not None and ptab.numel() != 13 * sum(plan.Ltaus):
        raise QfError(_lib.QF_ERR_INVALID, "preconditioner tables do not match the plan")
    dev = b.device
    b = _buf(b, "b", torch.complex128, plan.nsites, dev)
    phi32 = phi64.to(torch.float32) if phi32 is None else _buf(phi32, "phi32", torch.float32, phi64.numel(), dev)
    nbytes = _lib.lib().qf_pcg_mixed_ragged_workspace_bytes(plan.host.ctypes.data)
    work = _workspace(("ragged_mixed", tuple(plan.Ls), tuple(plan.Ltaus), dev), nbytes, dev)
    x = torch.empty_like(b) if x is None else _buf(x, "x", torch.complex128, b.numel(), dev)
    n = len(plan.Ls)
    err = torch.empty(n, dtype=torch.float64, device=dev)
    iters = torch.empty(n, dtype=torch.float64, device=dev)
    tab = ptab.to(torch.float32).contiguous() if ptab is not None else None
    st = _status_buf(status, n, dev)
    call("qf_pcg_mixed_solve_ragged", plan.host.ctypes.data, ptr(plan.dev), ptr(phi64), ptr(phi32), ptr(b), ptr(x), float(dtau),
         float(t), ptr(tab), float(rtol), int(max_iter), int(ru_period), ptr(work), work.numel(), ptr(err), ptr(iters), ptr(st),
         stream_handle())
    return x, iters, err
