"""CPU checks of the C-ABI library: it loads without a GPU, exports every symbol
the header declares, and its host-side functions (preconditioner table, CSR
verification, argument validation) behave."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "qed_fermion_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from qed_fermion_amd import _lib
    h = _lib.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(h, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    assert h.qf_version() == 2


def test_precon_table_matches_oracle():
    from qed_fermion_amd import _lib
    from oracle import oracle as O
    for Lt in (26, 60, 360):
        t = np.zeros((Lt, 13))
        _lib.call("qf_precon_table", Lt, t.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(t, O.precon_table(Lt))
    with pytest.raises(_lib.QfError) as e:
        _lib.call("qf_precon_table", 25, np.zeros((25, 13)).ctypes.data_as(ctypes.c_void_p))
    assert e.value.code == _lib.QF_ERR_LTAU


def test_precon_csr_roundtrip_and_rejection():
    import torch
    from qed_fermion_amd import precon
    from qed_fermion_amd._lib import QF_ERR_PRECON, QfError
    L, Lt = 4, 30
    d = precon.get_precon_man(L, Lt)
    csr = precon.coo_to_csr(d, device="cpu")
    tab = precon.table_from_csr(csr, L * L, Lt)
    ref = precon.band_table(Lt)
    # the CSR is complex64: values are the float32 constants, so the round trip is exact
    assert torch.equal(tab, ref)
    bad = d["values"].clone()
    bad[5] += 1.0          # one site differs from the others of its slice
    csr_bad = precon.coo_to_csr({"indices": d["indices"], "values": bad, "size": d["size"]}, device="cpu")
    with pytest.raises(QfError) as e:
        precon.table_from_csr(csr_bad, L * L, Lt)
    assert e.value.code == QF_ERR_PRECON


def test_argument_validation_without_gpu():
    from qed_fermion_amd import _lib
    h = _lib.lib()
    assert h.qf_pcg_workspace_bytes(1, 0, 1, 5, 10) == 0            # odd L rejected
    assert "even L" in h.qf_last_error().decode()
    assert h.qf_pcg_workspace_bytes(1, 0, 2, 36, 360) > 0
    rc = h.qf_apply(4, 0, None, 0, None, None, 1, 6, 10, 0.1, 1.0, None)
    assert rc == _lib.QF_ERR_INVALID
    rc = h.qf_apply(4, 0, None, 0, None, None, 1, 7, 10, 0.1, 1.0, None)
    assert rc == _lib.QF_ERR_ODD_L
    rc = h.qf_set_tuning(0, 3, 0, 0)
    assert rc == _lib.QF_ERR_INVALID
    out = (ctypes.c_int * 4)()
    assert h.qf_get_tiling(0, 1, 36, 360, out) == 0 and out[0] in (2, 4, 6)
    assert h.qf_get_tiling(0, 1, 66, 660, out) == 0 and out[0] == 6
    # right-hand sides per work-group of the plain M^dag M from phi: capped by the work-group
    # and the batch, one on band lattices; a shape entry keeps its tiling
    ns = ctypes.c_int()
    try:
        assert h.qf_set_tuning_shape(0, 36, 360, 32, 2, 23, 0) == 0
        assert h.qf_set_tuning_ns(0, 36, 360, 32, 7) == 0
        assert h.qf_get_rhs_per_group(0, 32, 36, 360, ctypes.byref(ns)) == 0 and ns.value == 3   # 3 x 324 lanes
        assert h.qf_get_tiling(0, 32, 36, 360, out) == 0 and (out[0], out[1]) == (2, 23)
        assert h.qf_set_tuning_ns(0, 36, 360, 2, 3) == 0
        assert h.qf_get_rhs_per_group(0, 2, 36, 360, ctypes.byref(ns)) == 0 and ns.value == 2    # the batch
        assert h.qf_set_tuning_ns(0, 70, 700, 8, 3) == 0
        assert h.qf_get_rhs_per_group(0, 8, 70, 700, ctypes.byref(ns)) == 0 and ns.value == 1    # bands
        assert h.qf_get_rhs_per_group(0, 32, 36, 100, ctypes.byref(ns)) == 0 and ns.value == 1   # untuned shape
        assert h.qf_set_tuning_ns(0, 36, 360, 32, -1) == _lib.QF_ERR_INVALID
    finally:
        h.qf_clear_tuning()
    # the allocation-free solver refuses a solve larger than its static scratch before any launch
    assert h.qf_pcg_noalloc_max_bytes() == 64 << 20
    rc = h.qf_pcg_solve_noalloc(1, None, 0, None, None, 64, 36, 360, 0.1, 1.0, None, 1e-9, 10, None, None, None, None)
    assert rc == _lib.QF_ERR_INVALID and "static scratch" in h.qf_last_error().decode()


def test_paired_x_switch():
    from qed_fermion_amd import _lib
    h = _lib.lib()
    assert h.qf_set_paired_x_updates(0) == 1          # on by default
    assert h.qf_set_paired_x_updates(5) == 0          # any nonzero value turns it on
    assert h.qf_set_paired_x_updates(1) == 1


def test_ragged_plan_layout():
    """qf_ragged_plan (host only): offsets, sizes and argument checks."""
    import ctypes
    import numpy as np
    from qed_fermion_amd import _lib
    lib = _lib.lib()
    Ls = np.array([12, 18, 24, 30], dtype=np.int32)
    Lts = np.array([120, 180, 240, 300], dtype=np.int32)
    n = len(Ls)
    nb = lib.qf_ragged_plan_bytes(n)
    assert nb > 0
    plan = np.zeros(nb, dtype=np.uint8)
    assert lib.qf_ragged_plan(n, Ls.ctypes.data, Lts.ctypes.data, plan.ctypes.data, nb) == 0
    head = np.frombuffer(plan[:48].tobytes(), dtype=np.int32)
    assert head[0] == n and head[1] == 2          # nprob, patch height
    assert head[2] == 256                          # 15 x 15 = 225 threads -> 4 waves
    nsites = np.frombuffer(plan[32:48].tobytes(), dtype=np.int64)
    assert nsites[0] == int(np.sum(Ls.astype(np.int64) ** 2 * Lts)) and nsites[1] == 13 * int(np.sum(Lts))
    # workspaces: the mixed solve adds x_lo, b_lo and the FP64 partials; complex128 adds the hop table
    nb32 = lib.qf_pcg_ragged_workspace_bytes(0, plan.ctypes.data)
    nbm = lib.qf_pcg_mixed_ragged_workspace_bytes(plan.ctypes.data)
    nb64 = lib.qf_pcg_ragged_workspace_bytes(1, plan.ctypes.data)
    assert 0 < nb32 < nbm and nb32 < nb64
    assert nbm - nb32 >= 2 * 8 * int(nsites[0])
    odd = np.array([12, 13], dtype=np.int32)
    assert lib.qf_ragged_plan(2, odd.ctypes.data, Lts.ctypes.data, plan.ctypes.data, nb) == _lib.QF_ERR_ODD_L
    assert lib.qf_ragged_plan(n, Ls.ctypes.data, Lts.ctypes.data, plan.ctypes.data, 8) == _lib.QF_ERR_INVALID


def test_buffer_validation_rejects_wrong_buffers():
    """Every caller-supplied buffer goes through one check (dtype, exact size, contiguity, then
    device) before its pointer is taken; a wrong buffer raises QfError instead of becoming an
    out-of-bounds access or a write into a temporary copy."""
    import torch
    from qed_fermion_amd import kernels as K
    from qed_fermion_amd._lib import QF_ERR_INVALID, QfError

    def rejects(fn, what):
        with pytest.raises(QfError) as e:
            fn()
        assert e.value.code == QF_ERR_INVALID and what in str(e.value), str(e.value)

    good = torch.zeros(8, dtype=torch.float64)
    rejects(lambda: K._buf(None, "x", torch.float64, 8), "required")
    rejects(lambda: K._buf(good.float(), "x", torch.float64, 8), "dtype")
    rejects(lambda: K._buf(good[:7], "x", torch.float64, 8), "elements")
    rejects(lambda: K._buf(torch.zeros(16, dtype=torch.float64)[::2], "x", torch.float64, 8), "contiguous")
    rejects(lambda: K._buf(good, "x", torch.float64, 8), "not on the GPU")
    rejects(lambda: K._vecs(torch.zeros(10, dtype=torch.complex64), 4), "multiple")
    rejects(lambda: K._vecs(torch.zeros(8, dtype=torch.float32), 4), "dtype")
    rejects(lambda: K._fields64(torch.zeros(2 * 4 * 4 * 3, dtype=torch.float32), 4, 3), "float64")
    # the public wrappers refuse CPU tensors: there is no CPU fallback
    rejects(lambda: K.apply("MdagM", torch.zeros(2 * 16 * 3), torch.zeros(16 * 3, dtype=torch.complex64), 4, 3),
            "GPU")
    rejects(lambda: K.boson_force(torch.zeros(2 * 16 * 3, dtype=torch.float64), 4, 3, 1.0, 1.0), "GPU")
