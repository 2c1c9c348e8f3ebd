"""ctypes binding of libqf_hip.so (the C ABI declared in include/qed_fermion_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
qed_fermion_amd/csrc``).  There is no fallback: if the shared object is
missing or a call fails, a :class:`QfError` is raised.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(_HERE, "libqf_hip.so")

QF_OK = 0
QF_ERR_INVALID = 1
QF_ERR_ODD_L = 2
QF_ERR_LTAU = 3
QF_ERR_HIP = 4
QF_ERR_UNSUPPORTED = 5
QF_ERR_PRECON = 6

QF_PREC_F32 = 0
QF_PREC_F64 = 1

QF_OP_M = 1
QF_OP_MDAG = 2
QF_OP_B = 3
QF_OP_MDAGM = 4

QF_SOLVE_CONVERGED = 0
QF_SOLVE_MAX_ITER = 1
QF_SOLVE_NONFINITE = 2
QF_SOLVE_BREAKDOWN = 3
QF_SOLVE_ABORTED = 4
SOLVE_STATUS_NAMES = {0: "converged", 1: "max_iter reached", 2: "non-finite residual or step (NaN/Inf in phi or b)",
                      3: "breakdown (Re(p^dag A p) <= 0)", 4: "aborted (a grid barrier of the persistent solve timed out)"}

QF_OBS_T1 = 0
QF_OBS_T21 = 1
QF_OBS_T2 = 2
QF_OBS_T4 = 3
QF_OBS_V = 4

_ERR_NAMES = {1: "QF_ERR_INVALID", 2: "QF_ERR_ODD_L", 3: "QF_ERR_LTAU", 4: "QF_ERR_HIP",
              5: "QF_ERR_UNSUPPORTED", 6: "QF_ERR_PRECON"}


class QfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


_P = ctypes.c_void_p
_I = ctypes.c_int
_LL = ctypes.c_longlong
_D = ctypes.c_double
_SZ = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/qed_fermion_amd.h
SIGNATURES = {
    "qf_version": (_I, []),
    "qf_last_error": (ctypes.c_char_p, []),
    "qf_set_tuning": (_I, [_I, _I, _I, _I]),
    "qf_set_tuning_shape": (_I, [_I, _I, _I, _I, _I, _I, _I]),
    "qf_clear_tuning": (_I, []),
    "qf_set_paired_x_updates": (_I, [_I]),
    "qf_get_tiling": (_I, [_I, _I, _I, _I, _P]),
    "qf_set_tuning_ns": (_I, [_I, _I, _I, _I, _I]),
    "qf_get_rhs_per_group": (_I, [_I, _I, _I, _I, _P]),
    "qf_apply": (_I, [_I, _I, _P, _LL, _P, _P, _I, _I, _I, _D, _D, _P]),
    "qf_hop_table": (_I, [_I, _P, _LL, _I, _I, _I, _D, _D, _P, _P]),
    "qf_apply_hop": (_I, [_I, _I, _P, _LL, _P, _P, _I, _I, _I, _D, _D, _P]),
    "qf_mhm_vec": (_I, [_I, _P, _P, _P, _I, _I, _I, _D, _P]),
    "qf_mh_vec": (_I, [_I, _P, _P, _P, _I, _I, _I, _D, _P]),
    "qf_precon_vec": (_I, [_I, _P, _P, _P, _I, _I, _I, _P]),
    "qf_precon_table": (_I, [_I, _P]),
    "qf_precon_table_from_csr": (_I, [_P, _P, _P, _P, _LL, _I, _I, _P]),
    "qf_pcg_workspace_bytes": (_SZ, [_I, _I, _I, _I, _I]),
    "qf_pcg_solve": (_I, [_I, _P, _LL, _P, _P, _I, _I, _I, _D, _D, _P, _D, _I, _P, _SZ, _P, _P, _P, _P]),
    "qf_pcg_noalloc_max_bytes": (_SZ, []),
    "qf_pcg_solve_noalloc": (_I, [_I, _P, _LL, _P, _P, _I, _I, _I, _D, _D, _P, _D, _I, _P, _P, _P, _P]),
    "qf_pcg_persistent_workspace_bytes": (_SZ, [_I, _I, _I, _I]),
    "qf_set_persistent_timeout_us": (_LL, [_LL]),
    "qf_pcg_solve_persistent": (_I, [_I, _P, _LL, _P, _P, _I, _I, _I, _D, _D, _P, _D, _I, _P, _SZ, _P, _P, _P, _P]),
    "qf_pcg_mixed_solve": (_I, [_P, _P, _LL, _P, _P, _I, _I, _I, _D, _D, _P, _D, _I, _I, _P, _SZ, _P, _P, _P, _P]),
    "qf_ragged_plan_bytes": (_SZ, [_I]),
    "qf_ragged_plan": (_I, [_I, _P, _P, _P, _SZ]),
    "qf_apply_ragged": (_I, [_I, _I, _P, _P, _P, _P, _P, _D, _D, _P]),
    "qf_pcg_ragged_workspace_bytes": (_SZ, [_I, _P]),
    "qf_pcg_solve_ragged": (_I, [_I, _P, _P, _P, _P, _P, _D, _D, _P, _D, _I, _P, _SZ, _P, _P, _P, _P]),
    "qf_pcg_ragged_persistent_workspace_bytes": (_SZ, [_I, _P]),
    "qf_pcg_solve_ragged_persistent": (_I, [_I, _P, _P, _P, _P, _P, _D, _D, _P, _D, _I, _P, _SZ, _P, _P, _P, _P]),
    "qf_pcg_mixed_ragged_workspace_bytes": (_SZ, [_P]),
    "qf_pcg_mixed_solve_ragged": (_I, [_P, _P, _P, _P, _P, _P, _D, _D, _P, _D, _I, _I, _P, _SZ, _P, _P, _P, _P]),
    "qf_fermion_force": (_I, [_I, _P, _LL, _P, _P, _I, _I, _I, _D, _D, _P]),
    "qf_boson_force": (_I, [_P, _P, _P, _I, _I, _I, _D, _D, _I, _P]),
    "qf_leapfrog_boson": (_I, [_P, _P, _P, _P, _P, _P, _D, _I, _I, _I, _I, _D, _D, _I, _P]),
    "qf_momentum_kick": (_I, [_P, _P, _P, _D, _P, _I, _I, _I, _P]),
    "qf_hamiltonian_workspace_bytes": (_SZ, [_I]),
    "qf_hamiltonian": (_I, [_I, _P, _P, _P, _P, _P, _P, _I, _I, _I, _D, _D, _I, _D, _P, _SZ, _P]),
    "qf_obs_phase": (_I, [_I, _P, _P, _LL, _P]),
    "qf_obs_fields": (_I, [_I, _I, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P]),
    "qf_obs_kspace": (_I, [_I, _P, _P, _LL, _I, _I, _I, _P, _I, _P]),
    "qf_obs_reduce": (_I, [_P, _I, _I, _P, _P]),
    "qf_obs_diag": (_I, [_I, _P, _P, _P, _I, _I, _I, _P, _I, _P]),
    "qf_obs_fused_grid": (_I, [_I, _I]),
    "qf_obs_fused": (_I, [_I, _I, _P, _P, _P, _P, _I, _I, _I, _I, _P, _I, _P]),
}

_lib = None


def lib():
    """Load libqf_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(SO_PATH):
            raise QfError(QF_ERR_HIP, f"{SO_PATH} is missing: run __graft_entry__.build() "
                                      "(make -C qed_fermion_amd/csrc) first; there is no CPU fallback")
        h = ctypes.CDLL(SO_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(rc):
    if rc != QF_OK:
        msg = lib().qf_last_error()
        raise QfError(rc, msg.decode() if msg else "")
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def stream_handle():
    """hipStream_t of torch's current stream on the current device."""
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())
