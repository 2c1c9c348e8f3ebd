"""Host logic of the autotuner (no GPU): the re-timing of the best candidates."""
from qed_fermion_amd import autotune


def test_refine_retimes_the_best_candidates_and_keeps_minima(monkeypatch):
    samples = iter([0.9, 1.2, 1.5, 1.0, 0.8, 1.6])      # rounds of (a, b, c), interleaved
    order = []
    monkeypatch.setattr(autotune, "_time", lambda fn, reps, warm=2: next(samples))
    table = {"a": 1.0, "b": 1.1, "c": 2.0, "d": 3.0}
    autotune._refine(table, order.append, lambda: None, reps=1, top=3, rounds=2)
    assert order == ["a", "b", "c", "a", "b", "c"]       # the slowest candidate is not re-timed
    assert table == {"a": 0.9, "b": 0.8, "c": 1.5, "d": 3.0}
    assert min(table, key=table.get) == "b"


def test_refine_scales_per_iteration_times(monkeypatch):
    monkeypatch.setattr(autotune, "_time", lambda fn, reps, warm=2: 20.0)
    table = {"x": 5.0, "y": 0.5}
    autotune._refine(table, lambda k: None, lambda: None, reps=1, scale=1.0 / 20, top=2, rounds=1)
    assert table == {"x": 1.0, "y": 0.5}


def test_valid_patch_heights_include_bands_where_a_slice_does_not_fit():
    f32, f64 = 0, 1
    assert autotune.valid_by(36, f32) == [2, 4, 6] and autotune.valid_by(36, f64) == [2]
    assert autotune.valid_by(66, f32) == [6, 2] and autotune.valid_by(66, f64) == [6, 2]   # 33 x 33 patches: bands
    assert autotune.valid_by(70, f32) == [2] and autotune.valid_by(96, f64) == [2]        # bands only
    assert autotune.valid_by(226, f32) == [2] and autotune.valid_by(228, f32) == []


def test_tau_run_candidates_fill_whole_rounds_of_work_groups():
    c = autotune._t_candidates(360, nrhs=32, ncu=256)
    assert 23 in c and 24 in c              # 32 x 16 runs = 512 work-groups: two per CU
    assert 45 in c                          # one per CU
    assert max(c) <= 48 and all(t <= 360 for t in c)
    assert autotune._t_candidates(10) == list(range(1, 11))


def test_rhs_per_group_candidates():
    assert autotune.ns_candidates(36, 2, 32) == [1, 2, 3]      # 324 lanes per slice, 1024 per work-group
    assert autotune.ns_candidates(36, 4, 32) == [1, 2, 3]      # 162 lanes, 512 per work-group
    assert autotune.ns_candidates(36, 6, 32) == [1, 2, 3, 4]   # 108 lanes; at most 4
    assert autotune.ns_candidates(36, 2, 2) == [1, 2]          # not more than the batch
    assert autotune.ns_candidates(36, 2, 1) == [1]
    assert autotune.ns_candidates(64, 2, 8) == [1]             # a slice fills the work-group
    assert autotune.ns_candidates(70, 2, 8) == [1]             # bands
