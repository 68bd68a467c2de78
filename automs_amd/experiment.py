"""One trial of the phase-map experiment (physical_exp_nni_model.py) as a
library call: the reference's 4321 scene geometry, the two channel sets of
build_channels (without and with the metasurface, the reference's own rays),
the four models and their optimisation, and the coverage statistics that the
reference's draw_optimization_result plots (go.py:944-964, :1004-1026).

  ap:  6 x 6 at 0.58 lambda, 0.3 m in front of wall 0's centre, 2.5 m above
       the floor, pointed by gen_direction_from_angle(0, 23.4352 - 90)
       (physical_exp_nni_model.py:36-57)
  Rx:  three layers of 84 x 75 over x in [-6, -1], y in [2.4, 6.5], 0.9 /
       1.05 / 1.2 m above the floor (:59-62)
  MS:  160 x 200 elements at lambda / 2 on walls[wall_id] at (0.5, ms_j),
       2 m above the floor, max phase 300 degrees (:64-74)
  models (:96-134): "AP no code book" (not optimised), "AP code book",
       "AP code book + MS not optimized" (metasurface phases frozen at 0),
       "AP code book + MS optimized"; 30 codes, go.py's capacity loss with
       w = 0.5, Adam lr 0.3 for `steps` steps keeping the best parameters.

run_trial returns the timings, each model's final and best loss (the trial
reports models[3]'s last loss to the tuner, :149) and the coverage
statistics: per Rx the best code's 20 log10 |y| clamped at -70 dB.
"""
from __future__ import annotations

import time
from typing import Dict, Optional

import torch

from . import channel, geometry
from .phasemap import Metasurface, PhaseMapModel, optimize_model

WAVE_LEN = 3e8 / 60e9
FLOOR_Z = -1.4161
# physical_exp_nni_model.py:7-10, the single-run placement
DEFAULT_PARAMS = {"ms_j": 0.11939808933765397, "wall_id": 1}
MODEL_NAMES = ("AP no code book", "AP code book", "AP code book + MS not optimized", "AP code book + MS optimized")


def trial_geometry(scene_folder: str, params: Optional[dict] = None, src_size=(6, 6), dst_size=(84, 75),
                   ms_size=(160, 200), heights=(0.9, 1.05, 1.2)):
    """AP direction, AP and Rx positions and the metasurface of one trial
    (physical_exp_nni_model.py:36-74)."""
    p = dict(DEFAULT_PARAMS, **(params or {}))
    wl = WAVE_LEN
    ap_dir = geometry.gen_direction_from_angle(0, 23.4352 - 90, [0, 0, 1], [0, 1, 0])
    walls = channel.get_walls(scene_folder)
    center_a, k_a, _ = geometry.gen_sth_from_wall(walls[0], (0.5, 0.5))
    center_a[2] = FLOOR_Z + 2.5
    src = geometry.gen_pos_matrix(center_a + k_a * 0.3, ap_dir, [0, 0, -1], src_size, wl * 0.58)
    dst = torch.cat([geometry.gen_pos_from_endpoint([-6, 2.4, dst_size[0]], [-1, 6.5, dst_size[1]], FLOOR_Z + h)
                     for h in heights], 0)
    c1, k1, i1 = geometry.gen_sth_from_wall(walls[int(p["wall_id"])], (0.5, float(p["ms_j"])))
    c1[2] = FLOOR_Z + 2.0
    pos, bh, n1 = geometry.gen_ms(c1, k1, i1, ms_size, wl / 2)
    ms = Metasurface(pos, ms_size, n1, n1, bh, 300)
    return ap_dir, src, dst, ms


@torch.no_grad()
def coverage_stats(model: PhaseMapModel, floor_db: float = -70.0) -> dict:
    """Per Rx the best code's received power 20 log10 |y| (go.py:164-174),
    NaN and values below floor_db set to floor_db (go.py:960-963); its mean,
    median and the fraction of Rx above the floor."""
    Y = model.received()
    amp = Y.abs().amax(0)
    db = 20 * torch.log10(amp)
    db = torch.where(torch.isnan(db), torch.full_like(db, floor_db), db).clamp(min=floor_db).double()
    return {"mean_db": float(db.mean()), "median_db": float(db.median()),
            "frac_above_floor": float((db > floor_db).double().mean()), "n_rx": int(db.numel())}


def build_trial_channels(scene_folder: str, src, dst, ms, ap_dir, max_reflection: int = 4) -> Dict[str, dict]:
    """The two channel sets of physical_exp_nni_model.py:76-90, traced at the
    same time (independent traces; the same channels as one after the
    other)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=2) as ex:
        noms = ex.submit(channel.build_channels, src, dst, [], ap_dir, scene_folder, WAVE_LEN, max_reflection)
        with_ms = ex.submit(channel.build_channels, src, dst, [ms], ap_dir, scene_folder, WAVE_LEN, max_reflection)
        return {"noms": noms.result(), "ms": with_ms.result()}


def trial_models(channels: Dict[str, dict], ms: Metasurface, code_num: int = 30, w: float = 0.5):
    """The four models of physical_exp_nni_model.py:96-134 (each with its own
    random AP phases, the metasurface phases starting at 0)."""
    m = [PhaseMapModel(MODEL_NAMES[0], channels["noms"], [], code_num=code_num, w=w, objective="capacity", seed=0),
         PhaseMapModel(MODEL_NAMES[1], channels["noms"], [], code_num=code_num, w=w, objective="capacity", seed=1),
         PhaseMapModel(MODEL_NAMES[2], channels["ms"], [ms], code_num=code_num, w=w, objective="capacity", seed=2),
         PhaseMapModel(MODEL_NAMES[3], channels["ms"], [ms], code_num=code_num, w=w, objective="capacity", seed=3)]
    m[2].ms_theta[0].requires_grad_(False)  # physical_exp_nni_model.py:136
    return m


def run_trial(scene_folder: str, params: Optional[dict] = None, steps: int = 200, code_num: int = 30, w: float = 0.5,
              seed: int = 5, max_reflection: int = 4, ms_size=(160, 200), dst_size=(84, 75)) -> dict:
    """physical_exp_nni_model.py end to end; returns timings (seconds of the
    four traces and of the three optimisations, synchronised), the models'
    losses and coverage statistics, and the models themselves."""
    channel.init(seed=seed)  # go.py:530: the capacity loss's row masks come from torch's CPU generator
    ap_dir, src, dst, ms = trial_geometry(scene_folder, params, ms_size=ms_size, dst_size=dst_size)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ch = build_trial_channels(scene_folder, src, dst, ms, ap_dir, max_reflection)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    models = trial_models(ch, ms, code_num, w)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for m in models[1:]:
        # one HIP-graph launch per step; the capacity loss's row masks are the
        # eager loop's (drawn in the same order from the CPU generator)
        optimize_model(m, lr=0.3, step=steps, graph=True)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out = {"params": dict(DEFAULT_PARAMS, **(params or {})), "steps": steps,
           "seconds_traces": t1 - t0, "seconds_models": t2 - t1, "seconds_optimise": t3 - t2,
           "seconds_total": t3 - t0, "shapes": {k: {kk: list(v.shape) for kk, v in c.items()} for k, c in ch.items()},
           "models": {}}
    for m in models:
        loss = m.loss_list[-1] if m.loss_list else None
        with torch.no_grad():
            final = float(m().item())
        out["models"][m.name] = {"last_loss": loss, "best_loss": min(m.loss_list) if m.loss_list else None,
                                 "loss_at_best_params": final, **coverage_stats(m)}
    out["final_result"] = out["models"][MODEL_NAMES[3]]["last_loss"]  # :149 reports this to the tuner
    out["_models"], out["_channels"] = models, ch
    return out


@torch.no_grad()
def coverage_with_phases(channels: dict, ms_list, ap_theta: torch.Tensor, ms_theta=None, code_num: int = 30) -> dict:
    """Coverage statistics of given AP codebook phases (and metasurface phases)
    on these channels, e.g. the reference run's optimised phases on this
    tracer's channels."""
    m = PhaseMapModel("given", channels, ms_list, code_num=code_num, objective="capacity",
                      antenna_theta=torch.as_tensor(ap_theta, dtype=torch.float32),
                      ms_theta=None if ms_theta is None else [torch.as_tensor(ms_theta, dtype=torch.float32)])
    return coverage_stats(m)


def mean_gain_db(a: dict, b: dict) -> float:
    return a["mean_db"] - b["mean_db"]


__all__ = ["DEFAULT_PARAMS", "MODEL_NAMES", "trial_geometry", "build_trial_channels", "trial_models", "run_trial",
           "coverage_stats", "coverage_with_phases", "mean_gain_db"]
