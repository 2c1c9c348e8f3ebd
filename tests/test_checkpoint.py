"""On-disk formats: run tags and file names, boson ensembles, run checkpoints, the P cache.

The expected run tags are the checkpoint names the reference's own analysis
scripts open (known answers for hmc_sampler_batch.py:331's format and the Nrv
rule at 58-62).  Those runs predate the ``max_iter`` field, which the current
tag carries between cg_rtol and max_block_idx, so the comparison takes it out.
"""
import collections
import math
import os
import pickle

import pytest
import torch

from qed_fermion_amd import checkpoint, kernels, precon
from qed_fermion_amd.hmc_sampler import HmcSampler

pytestmark = pytest.mark.filterwarnings("ignore:Sparse CSR tensor support")

# (Lx, bs, K, Nstep, compact, file name) from
# observables/spin_r_cmp_large7_spsm/plot_fit_spsm_r.py:97-101 and
# observables/spin_r_noncmpK0_large1_spsm_sup/plot_fit_spsm_r.py:94-98
REFERENCE_CKPT_NAMES = [
    (46, 2, 1.0, 10000, True, "ckpt_N_hmc_46_Ltau_460_Nstp_10000_bs2_Jtau_1.2_K_1_dtau_0.1_delta_0.028_N_leapfrog_5_m_1_cg_rtol_1e-09_max_block_idx_1_gear0_steps_1000_dt_deque_max_len_5_Nrv_40_cmp_True_step_10000.pt"),
    (56, 1, 1.0, 10000, True, "ckpt_N_hmc_56_Ltau_560_Nstp_10000_bs1_Jtau_1.2_K_1_dtau_0.1_delta_0.028_N_leapfrog_5_m_1_cg_rtol_1e-09_max_block_idx_1_gear0_steps_1000_dt_deque_max_len_5_Nrv_30_cmp_True_step_10000.pt"),
    (60, 1, 1.0, 6800, True, "ckpt_N_hmc_60_Ltau_600_Nstp_6800_bs1_Jtau_1.2_K_1_dtau_0.1_delta_0.028_N_leapfrog_5_m_1_cg_rtol_1e-09_max_block_idx_1_gear0_steps_1000_dt_deque_max_len_5_Nrv_30_cmp_True_step_6800.pt"),
    (46, 2, 0.0, 10000, False, "ckpt_N_hmc_46_Ltau_460_Nstp_10000_bs2_Jtau_1.2_K_0_dtau_0.1_delta_0.028_N_leapfrog_5_m_1_cg_rtol_1e-09_max_block_idx_1_gear0_steps_1000_dt_deque_max_len_5_Nrv_40_cmp_False_step_10000.pt"),
    (56, 1, 0.0, 10000, False, "ckpt_N_hmc_56_Ltau_560_Nstp_10000_bs1_Jtau_1.2_K_0_dtau_0.1_delta_0.028_N_leapfrog_5_m_1_cg_rtol_1e-09_max_block_idx_1_gear0_steps_1000_dt_deque_max_len_5_Nrv_30_cmp_False_step_10000.pt"),
]


@pytest.fixture(autouse=True)
def _clean_env(monkeypatch):
    for k in ("gear0_steps", "Nrv", "compact"):
        monkeypatch.delenv(k, raising=False)


class _Tag:
    """The attributes the run tag reads, without building a sampler (lattices up to L = 60)."""

    def __init__(self, Lx, bs, K, Nstep, compact, J=1.2, dtau=0.1, cg_rtol=1e-9, max_iter=400):
        self.Lx, self.Ly, self.Ltau, self.bs, self.N_step = Lx, Lx, 10 * Lx, bs, Nstep
        self.Nf, self.dtau = 2, dtau
        self.J = J / dtau * self.Nf / 4          # stored as in hmc_sampler_batch.py:169-175
        self.K = K * dtau * self.Nf / 2
        self.N_leapfrog, self.m, self.cg_rtol, self.max_iter = 5, 1, cg_rtol, max_iter
        self.max_tau_block_idx = 1
        self.threshold_queue = [collections.deque(maxlen=5)]
        self.compact = compact


@pytest.mark.parametrize("Lx,bs,K,Nstep,compact,name", REFERENCE_CKPT_NAMES)
def test_run_tag_matches_reference_files(Lx, bs, K, Nstep, compact, name):
    t = _Tag(Lx, bs, K, Nstep, compact)
    got = checkpoint.checkpoint_file_name(t, Nstep) + ".pt"
    assert got.replace("_max_iter_400", "") == name


def test_default_nrv_thresholds():
    # boundaries of hmc_sampler_batch.py:58-60 (sites = Lx*Lx*Ltau)
    assert checkpoint.default_nrv(6, 10) == 30          # below 80000
    assert checkpoint.default_nrv(20, 200) == 40        # 80000
    assert checkpoint.default_nrv(50, 500) == 40        # 1250000
    assert checkpoint.default_nrv(52, 520) == 30
    assert checkpoint.default_nrv(60, 600) == 30        # 2160000
    assert checkpoint.default_nrv(62, 620) == 20


def test_sampler_tag_is_the_run_tag():
    s = HmcSampler(Lx=6, Ltau=60, bs=2, J=1.2, K=1.0, Nstep=10000, device="cpu", max_iter=400)
    assert s.specifics == checkpoint.specifics(s)
    assert s.specifics.startswith("hmc_6_Ltau_60_Nstp_10000_bs2_Jtau_1.2_K_1_dtau_0.1_delta_0.028_")


def test_ensemble_file_name_rules():
    # cg_rtol 1e-15 -> 1e-09, max_iter dropped, Nstp -> 10000 -> 6800 at Lx = 60,
    # cmp_True -> cmp_False, Jtau -> 1 at Lx = 60 (hmc_sampler_batch.py:981-991)
    t = _Tag(60, 1, 1.0, 5000, True, J=1.2, cg_rtol=1e-15, max_iter=123)
    assert checkpoint.ensemble_file_name(t, 4000) == (
        "boson_hmc_60_Ltau_600_Nstp_6800_bs1_Jtau_1_K_1_dtau_0.1_delta_0.028_N_leapfrog_5_m_1_cg_rtol_1e-09"
        "_max_block_idx_1_gear0_steps_1000_dt_deque_max_len_5_Nrv_30_cmp_False_4000.pt")
    t = _Tag(12, 2, 1.0, 3000, False, J=0.5)
    assert checkpoint.ensemble_file_name(t, 4500) == (
        "boson_hmc_12_Ltau_120_Nstp_10000_bs2_Jtau_0.5_K_1_dtau_0.1_delta_0.028_N_leapfrog_5_m_1_cg_rtol_1e-09"
        "_max_block_idx_1_gear0_steps_1000_dt_deque_max_len_5_Nrv_30_cmp_False_4500.pt")


def _sampler(root=None, **kw):
    return HmcSampler(Lx=6, Ltau=60, bs=2, device="cpu", ckpt_root=root, **kw)


def test_boson_ensemble_round_trip(tmp_path):
    a = _sampler()
    saved = {}
    for step in (4000, 4500, 6000):
        a.initialize_boson_uniform()
        a.delta_t_tensor = torch.full((2,), 0.01 * step / 1000)
        path = checkpoint.save_boson_ensemble(a, str(tmp_path), step)
        assert os.path.basename(path) == checkpoint.ensemble_file_name(a, step)
        saved[step] = (a.boson.clone(), a.delta_t_tensor.clone())
    d = torch.load(path, weights_only=True)
    assert set(d) == {"boson", "delta_t_tensor"} and d["boson"].shape == (2, 2, 6, 6, 60)

    b = _sampler(str(tmp_path))          # reset() starts from the ensemble
    assert any(torch.equal(b.boson, x) for x, _ in saved.values())
    assert any(torch.equal(b.delta_t_tensor, dt) for _, dt in saved.values())
    bosons, dts = checkpoint.load_boson_ensemble(b, str(tmp_path))
    assert len(bosons) == 3 and len(dts) == 3


def test_boson_ensemble_bare_tensor_and_fallback(tmp_path):
    a = _sampler()
    folder = tmp_path / "check_points" / "boson_ensemble"
    folder.mkdir(parents=True)
    bare = torch.rand(2, 2, 6, 6, 60)
    torch.save(bare, folder / checkpoint.ensemble_file_name(a, 5000))
    b = _sampler(str(tmp_path))
    assert torch.equal(b.boson, bare)
    assert torch.allclose(b.delta_t_tensor, torch.full((2,), 0.06))   # no delta_t_tensor in a bare file
    empty = tmp_path / "none"
    c = _sampler(str(empty))
    assert checkpoint.initialize_boson_ensemble(c, str(empty)) == 0
    assert c.boson.shape == (2, 2, 6, 6, 60)


def test_checkpoint_resume(tmp_path):
    a = _sampler()
    a.initialize_boson_uniform()
    a.step, a.cur_step = 120, 120
    lists = {"cg_r_err_list": torch.rand(200, 2), "delta_t_list": torch.rand(200, 2)}
    path = checkpoint.save_checkpoint(a, str(tmp_path), suffix="bench", lists=lists)
    assert os.path.basename(path) == f"ckpt_N_{a.specifics}_step_120.pt"
    assert os.path.basename(os.path.dirname(path)) == "hmc_check_point_bench"
    b = _sampler()
    res = checkpoint.load_checkpoint(path, b)
    assert torch.equal(b.boson, a.boson) and b.step == 120 and b.cur_step == 120
    assert torch.equal(res["cg_r_err_list"], lists["cg_r_err_list"])
    with pytest.raises(ValueError):
        checkpoint.load_checkpoint(path, HmcSampler(Lx=8, Ltau=60, bs=2, device="cpu"))


def test_precon_cache_written_and_read(tmp_path):
    s = _sampler(str(tmp_path))
    s.reset_precon()
    path = tmp_path / "preconditioners" / "precon_ckpt_L_6_Ltau_60_dtau_0.1_t_1.pt"
    assert path.exists()
    d = torch.load(path, weights_only=True)
    assert set(d) == {"indices", "values", "size"}
    assert d["values"].dtype == torch.complex64 and tuple(d["size"]) == (2160, 2160)
    ref = precon.get_precon_man(6, 60)
    assert torch.equal(d["indices"], ref["indices"]) and torch.equal(d["values"], ref["values"])
    # the table read back is the manual band table at the file's complex64 precision
    want = precon.band_table(60).to(torch.float32).to(torch.float64)
    assert torch.equal(s.precon_table.cpu(), want)
    # a second sampler reads the cached file
    t = _sampler(str(tmp_path))
    t.reset_precon()
    assert torch.equal(t.precon_table, s.precon_table)


def test_precon_cache_structure_checked(tmp_path):
    d = precon.get_precon_man(6, 60)
    idx = torch.cat([d["indices"], torch.tensor([[0], [7]])], dim=1)     # one entry off the tau band
    val = torch.cat([d["values"], torch.tensor([0.5 + 0j], dtype=torch.complex64)])
    path = checkpoint.save_precon({"indices": idx, "values": val, "size": d["size"]}, str(tmp_path / "p.pt"))
    with pytest.raises(kernels.QfError):
        checkpoint.precon_table_from_file(path, 6, 60)
    with pytest.raises(ValueError):
        checkpoint.precon_table_from_file(path, 8, 60)       # size does not match the lattice


class _NotData:
    pass


def test_loads_execute_nothing(tmp_path):
    path = str(tmp_path / "obj.pt")
    torch.save({"indices": torch.zeros(2, 1, dtype=torch.int64), "values": torch.zeros(1), "size": (1, 1),
                "extra": _NotData()}, path)
    with pytest.raises(pickle.UnpicklingError):
        checkpoint.load_precon(path)
    with pytest.raises(pickle.UnpicklingError):
        checkpoint.load_checkpoint(path)


def test_pi_flux_start_matches_reference_recipe():
    # hmc_sampler_batch.py:1026-1043: the last chain is the staggered pi flux, the others add
    # tau-independent noise of standard deviation 0.02 ... 0.08
    torch.manual_seed(3)
    s = HmcSampler(Lx=8, Ltau=40, bs=4, device="cpu")
    b = s.boson
    sign = torch.tensor([[(-1) ** (x + y) for y in range(8)] for x in range(8)], dtype=b.dtype)
    stag = torch.stack([sign * math.pi / 4, -sign * math.pi / 4]).unsqueeze(-1).expand(2, 8, 8, 40)
    assert torch.equal(b[-1], stag)
    dev = b[:-1] - stag
    assert torch.equal(dev, dev[..., :1].expand_as(dev))           # constant along tau
    sd = dev[..., 0].reshape(3, -1).std(dim=1)
    assert (sd[0] < sd[1] < sd[2]) and 0.01 < sd[0] < 0.04 and 0.05 < sd[2] < 0.12
