"""ctypes front end of the FP64 CPU oracle (oracle/qf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker and by ``bench.py``'s cpu_baseline leg.  The product package
``qed_fermion_amd`` never imports this module.

Arrays are numpy, in the reference's permuted layout (hmc_sampler_batch.py:2166):
``phi`` float64 ``[Nt, L, L, 2]`` (tau, y, x, mu) and fermion vectors complex128
``[Nt, L, L]`` (tau, y, x), flattened freely.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libqf_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
        i, d, v = ctypes.c_int, ctypes.c_double, None
        sig = {
            "qfo_apply_B": (v, [_D, _D, _D, i, d, d]),
            "qfo_apply_Bpartial": (v, [_D, _D, _D, i, d, d, i]),
            "qfo_apply_M": (v, [_D, _D, _D, i, i, d, d]),
            "qfo_apply_Mdag": (v, [_D, _D, _D, i, i, d, d]),
            "qfo_apply_MdagM": (v, [_D, _D, _D, i, i, d, d]),
            "qfo_precon_table": (i, [i, _D]),
            "qfo_apply_P": (v, [_D, _D, _D, i, i]),
            "qfo_pcg": (i, [_D, _D, _D, i, i, d, d, _D, d, i, i, _D]),
            "qfo_force": (v, [_D, _D, _D, i, i, d, d]),
            "qfo_action_tau": (d, [_D, i, i, d, i]),
            "qfo_action_plaq": (d, [_D, i, i, d]),
            "qfo_force_b_tau": (v, [_D, _D, i, i, d, i]),
            "qfo_force_b_plaq": (v, [_D, _D, i, i, d]),
            "qfo_trajectory": (v, [_D, _D, _D, i, i, d, d, d, d, i, d, d, i, _D, d, i, i, _D, _D, _D]),
            "qfo_num_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def couplings(J=1.0, K=1.0, dtau=0.1, Nf=2):
    """(cJ, cK): tau-action coefficient 1/(J' dtau^2) with J' = J/dtau*Nf/4 and
    plaquette coupling K' = K*dtau*Nf/2 (hmc_sampler_batch.py:173-174, 1344)."""
    Jp = J / dtau * Nf / 4
    return 1.0 / (Jp * dtau ** 2), K * dtau * Nf / 2


def apply_B(phi_s, v, L, dtau=0.1, t=1.0, nfac=7):
    phi_s, v = _f64(phi_s), _c128(v)
    out = np.empty_like(v)
    lib().qfo_apply_Bpartial(_p(phi_s), _p(v), _p(out), L, dtau, t, nfac)
    return out


def _op(name, phi, x, L, Lt, dtau, t):
    phi, x = _f64(phi), _c128(x)
    y = np.empty_like(x)
    getattr(lib(), name)(_p(phi), _p(x), _p(y), L, Lt, dtau, t)
    return y


def apply_M(phi, x, L, Lt, dtau=0.1, t=1.0):
    return _op("qfo_apply_M", phi, x, L, Lt, dtau, t)


def apply_Mdag(phi, x, L, Lt, dtau=0.1, t=1.0):
    return _op("qfo_apply_Mdag", phi, x, L, Lt, dtau, t)


def apply_MdagM(phi, x, L, Lt, dtau=0.1, t=1.0):
    return _op("qfo_apply_MdagM", phi, x, L, Lt, dtau, t)


def precon_table(Lt):
    tab = np.zeros((Lt, 13), dtype=np.float64)
    if lib().qfo_precon_table(Lt, _p(tab)) != 0:
        raise ValueError("preconditioner needs Ltau >= 26")
    return tab


def apply_P(table, r, Vs, Lt):
    table, r = _f64(table), _c128(r)
    z = np.empty_like(r)
    lib().qfo_apply_P(_p(table), _p(r), _p(z), Vs, Lt)
    return z


def pcg(phi, b, L, Lt, dtau=0.1, t=1.0, table=None, rtol=1e-9, max_iter=400, early_exit=True):
    """Returns (x, iters, err) with the reference's iteration counter."""
    phi, b = _f64(phi), _c128(b)
    x = np.empty_like(b)
    err = ctypes.c_double(0.0)
    tp = None
    if table is not None:
        table = _f64(table)
        tp = _p(table)
    it = lib().qfo_pcg(_p(phi), _p(b), _p(x), L, Lt, dtau, t, tp, rtol, max_iter, 1 if early_exit else 0,
                       ctypes.byref(err))
    return x, it, err.value


def force(phi, xi, L, Lt, dtau=0.1, t=1.0):
    phi, xi = _f64(phi), _c128(xi)
    F = np.empty_like(phi)
    lib().qfo_force(_p(phi), _p(xi), _p(F), L, Lt, dtau, t)
    return F


def action_tau(phi, L, Lt, cJ, compact=False):
    return lib().qfo_action_tau(_p(_f64(phi)), L, Lt, cJ, int(bool(compact)))


def action_plaq(phi, L, Lt, cK):
    return lib().qfo_action_plaq(_p(_f64(phi)), L, Lt, cK)


def force_b_tau(phi, L, Lt, cJ, compact=False):
    phi = _f64(phi)
    F = np.empty_like(phi)
    lib().qfo_force_b_tau(_p(phi), _p(F), L, Lt, cJ, int(bool(compact)))
    return F


def force_b_plaq(phi, L, Lt, cK):
    phi = _f64(phi)
    F = np.empty_like(phi)
    lib().qfo_force_b_plaq(_p(phi), _p(F), L, Lt, cK)
    return F


def trajectory(phi, p, R, L, Lt, dtau=0.1, t=1.0, J=1.0, K=1.0, compact=False, m=1.0, dt=0.06,
               n_leap=5, table=None, rtol=1e-9, max_iter=400, early_exit=True):
    """One leapfrog_proposer5 trajectory.  Returns (phi_new, p_new, H0, H_fin, err)."""
    cJ, cK = couplings(J, K, dtau)
    phi = _f64(phi).copy()
    p = _f64(p).copy()
    R = _c128(R)
    H0, Hf, err = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    tp = None
    if table is not None:
        table = _f64(table)
        tp = _p(table)
    lib().qfo_trajectory(_p(phi), _p(p), _p(R), L, Lt, dtau, t, cJ, cK, int(bool(compact)), m, dt, n_leap,
                         tp, rtol, max_iter, 1 if early_exit else 0,
                         ctypes.byref(H0), ctypes.byref(Hf), ctypes.byref(err))
    return phi, p, H0.value, Hf.value, err.value


def num_threads():
    return lib().qfo_num_threads()
