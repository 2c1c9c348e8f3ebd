"""adjust_delta_t (hmc_sampler_batch.py:699-724) on the CPU: the step-size rule against a
loop written from the reference's statement of it, on random threshold histories."""
import collections
import random

import torch

from qed_fermion_amd.trajectory import adjust_delta_t


def _reference_rule(dt, queues, lower, upper):
    # the reference's per-chain loop: skip partial deques; x0.9 below the window, x1.1 above,
    # clear the deque only when dt changed
    out = list(dt)
    for b in range(len(out)):
        if len(queues[b]) < queues[b].maxlen:
            continue
        avg = sum(queues[b]) / len(queues[b])
        if avg < lower:
            out[b] = out[b] * 0.9
        elif upper < avg:
            out[b] = out[b] * 1.1
        else:
            continue
        queues[b].clear()
    return out


def test_adjust_delta_t_matches_reference_rule():
    rng = random.Random(7)
    bs, maxlen = 6, 5
    dt = torch.full((bs,), 0.06, dtype=torch.float64)
    ref_dt = [0.06] * bs
    qa = [collections.deque(maxlen=maxlen) for _ in range(bs)]
    qb = [collections.deque(maxlen=maxlen) for _ in range(bs)]
    for step in range(200):
        for b in range(bs):
            th = min(1.0, rng.random() * (0.3 + 0.2 * b))   # chains drift to different windows
            qa[b].append(th)
            qb[b].append(th)
        adjust_delta_t(dt, qa, 0.25, 0.5)
        ref_dt = _reference_rule(ref_dt, qb, 0.25, 0.5)
        assert [len(q) for q in qa] == [len(q) for q in qb], step
        assert torch.allclose(dt, torch.tensor(ref_dt, dtype=torch.float64), rtol=0, atol=0), step


def test_adjust_delta_t_window_edges():
    dt = torch.tensor([1.0, 1.0, 1.0, 1.0], dtype=torch.float64)
    qs = [collections.deque([0.2] * 5, maxlen=5), collections.deque([0.6] * 5, maxlen=5),
          collections.deque([0.25] * 5, maxlen=5), collections.deque([0.1] * 4, maxlen=5)]
    changed = adjust_delta_t(dt, qs)
    assert changed == [0, 1]
    assert dt.tolist() == [0.9, 1.1, 1.0, 1.0]      # 0.25 is inside the window; a partial deque waits
    assert [len(q) for q in qs] == [0, 0, 5, 4]      # adjusted deques are cleared
