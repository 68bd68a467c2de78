"""The placement search (hyper-config tuner, physical_exp_nni_main.py:12-29
with the trial of physical_exp_nni_model.py:96-149 as its objective): plain
simulated annealing, every candidate judged by a traced and optimised
trial, against annealing steered by the committed surrogate network, whose
accepted candidates the tracer re-scores.  Short searches at reduced rays:
both return a placement whose value is its judged trial score, the steered
search calls the judge only for accepted candidates, and the reported
values are reproducible by judging the returned placements again."""
from pathlib import Path

import numpy as np
import pytest

from _parity import flush

pytestmark = pytest.mark.gpu
DATA = Path(__file__).resolve().parents[1] / "automs_amd" / "data"


def test_anneal_plain_vs_surrogate(gpu):
    import torch

    from automs_amd import surrogate as S
    flush("case: placement annealing, plain vs surrogate-steered (scene 1000, 2^16 rays, 10-step trials)")
    net = S.SurrogateNet()
    net.load(str(DATA / "surrogate_room_trial.npy"))
    occ = S.occupancy(1000)
    pt = S.PlacementTracer(1000, 1 << 16, 1 << 16, trial_steps=10)
    steps = 40
    t0, t1 = 0.05, 0.0005
    pt.n_traces = 0
    b1, v1, a1 = S.anneal(pt.trial_score, steps, seed=1, t0=t0, t1=t1)
    n1 = pt.n_traces
    pt.n_traces = 0
    b2, v2, a2 = S.anneal(lambda p: float(S.predict_score(net, occ, [p])[0]), steps, seed=1, t0=t0, t1=t1,
                          rescore=pt.trial_score)
    n2 = pt.n_traces
    flush(f"  plain: best {v1:.4f} dB, {n1} trials, {a1} accepted; steered: best {v2:.4f} dB, {n2} trials, "
          f"{a2} accepted")
    assert n1 == steps + 1  # the start and every candidate
    assert n2 == a2 + 1 <= steps + 1  # the start and every accepted candidate
    for b, v in ((b1, v1), (b2, v2)):
        again = pt.trial_score(b)
        flush(f"  placement {b.__dict__}: reported {v:.4f} dB, judged again {again:.4f} dB")
        assert np.isfinite(v) and abs(again - v) < 0.02
    pt.ts.close()
    torch.cuda.empty_cache()
