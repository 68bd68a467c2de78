"""The reference's own phase-map run (physical_exp_nni_model.py: the 4321
scene, 6 x 6 AP, 3 x 84 x 75 Rx, a 160 x 200 metasurface on wall 1, 4
reflections, the reference's TEA/LCG rays, three models of 200 Adam steps),
end to end through automs_amd.experiment, against the statistics of the
reference's committed result (tests/golden/reference_run, written by
tests/golden/make_reference_run_stats.py from output/0_STANDALONE).

The reference's channels are Monte-Carlo traces on an NVIDIA GPU and its AP
start phases come from the CUDA generator, so its exact numbers are parity
unpinned; the bars are loose sanity bars: every model's mean per-Rx coverage
(best code, -70 dB floor) within 1 dB of the reference's, the fraction of Rx
above the floor within 0.03, the gain of the optimised metasurface over the
codebook alone within 0.5 dB of the reference's, and the reference's own
optimised phases on these channels within 0.5 dB of the coverage they gave
it."""
import json
from pathlib import Path

import numpy as np
import pytest

from _parity import flush

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def test_reference_run_end_to_end(gpu):
    import torch

    from automs_amd import experiment as E
    folder = str(GOLD / "scene_4321")
    flush("case: reference run (4321 scene, 160 x 200 metasurface, 3 x 84 x 75 Rx, 3 models x 200 steps)")
    r = E.run_trial(folder)
    ref = json.load(open(GOLD / "reference_run" / "stats.json"))
    flush(f"  {r['seconds_total']:.2f} s ({r['seconds_traces']:.2f} s traces, {r['seconds_optimise']:.2f} s "
          f"optimisation); shapes {r['shapes']}")
    assert r["shapes"]["ms"]["ms{0}_to_dst"] == [32000, 18900]
    assert r["shapes"]["noms"]["src_to_dst"] == [36, 18900]
    for name in E.MODEL_NAMES:
        m, q = r["models"][name], ref[name]
        flush(f"  {name}: mean {m['mean_db']:.2f} dB (reference {q['mean_db']:.2f}), above -70 dB "
              f"{m['frac_above_floor']:.3f} ({q['frac_above_floor']:.3f})")
        assert abs(m["mean_db"] - q["mean_db"]) < 1.0
        assert abs(m["frac_above_floor"] - q["frac_above_floor"]) < 0.03
    gain = r["models"][E.MODEL_NAMES[3]]["mean_db"] - r["models"][E.MODEL_NAMES[1]]["mean_db"]
    gain_ref = ref[E.MODEL_NAMES[3]]["mean_db"] - ref[E.MODEL_NAMES[1]]["mean_db"]
    flush(f"  gain of the optimised metasurface over the codebook alone: {gain:.2f} dB (reference {gain_ref:.2f})")
    assert gain > 0.5 and abs(gain - gain_ref) < 0.5
    p = np.load(GOLD / "reference_run" / "params.npz")
    _, _, _, ms = E.trial_geometry(folder)
    ch = r["_channels"]
    cb = E.coverage_with_phases(ch["noms"], [], p["ap_noms"])
    full = E.coverage_with_phases(ch["ms"], [ms], p["ap_ms"], p["ms_theta0"].reshape(1, -1))
    flush(f"  the reference's optimised phases on these channels: codebook {cb['mean_db']:.2f} dB "
          f"(reference {ref[E.MODEL_NAMES[1]]['mean_db']:.2f}), codebook + metasurface {full['mean_db']:.2f} dB "
          f"(reference {ref[E.MODEL_NAMES[3]]['mean_db']:.2f})")
    assert abs(cb["mean_db"] - ref[E.MODEL_NAMES[1]]["mean_db"]) < 0.5
    assert abs(full["mean_db"] - ref[E.MODEL_NAMES[3]]["mean_db"]) < 0.5
    del r, ch
    torch.cuda.empty_cache()
