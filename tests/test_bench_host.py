"""Host logic of bench.py (no GPU): the roofline's traffic figure comes from the counter
measurement of the tiling the run used, and the one-GPU contract."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_traffic_follows_the_tiling():
    with open(os.path.join(ROOT, "profiles", "mdagm_traffic.json")) as f:
        rec = json.load(f)
    seen = set()
    for e in rec["entries"]:
        key = (e["kind"], e["nrhs"], e["tiling"]["by"], e["tiling"]["t"], e["tiling"].get("ns", 1))
        assert key not in seen                      # one entry per shape and tiling
        seen.add(key)
        got = bench.traffic_of(e["L"], e["Ltau"], e["nrhs"], e["kind"], dict(e["tiling"]))
        assert got == e["traffic_bytes"]
        # at least the algorithmic bytes; the excess is the tau halo, which shrinks with the run length
        assert e["algorithmic_bytes"] <= e["traffic_bytes"] < 1.25 * e["algorithmic_bytes"]
    for kind in ("per_chain", "per_chain_phi"):
        for by in (2, 4):
            runs = sorted((e["tiling"]["t"], e["traffic_bytes"]) for e in rec["entries"]
                          if e["kind"] == kind and e["tiling"]["by"] == by and e["tiling"].get("ns", 1) == 1)
            assert [b for _, b in runs] == sorted((b for _, b in runs), reverse=True)   # longer runs, less halo
    # from phi: (24 + 8 * 3 / t) bytes per site per chain, the tau halo of a t-slice run
    for e in rec["entries"]:
        if e["kind"] == "per_chain_phi":
            t = e["tiling"]["t"]
            assert abs(e["traffic_over_algorithmic"] - (24 + 24 / t) / 24) < 0.01, e["tiling"]
    assert bench.traffic_of(36, 360, 32, "per_chain_phi", {"by": 2, "t": 24}) is not None
    assert bench.traffic_of(36, 360, 32, "per_chain_phi", {"by": 2, "t": 24, "ns": 1}) is not None
    assert bench.traffic_of(36, 360, 32, "per_chain_phi", {"by": 2, "t": 7, "ns": 2}) is None   # not measured
    assert bench.traffic_of(36, 360, 32, "per_chain", {"by": 6, "t": 7}) is None      # not measured
    assert bench.traffic_of(36, 360, 31, "per_chain", {"by": 4, "t": 12}) is None     # other shape


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
