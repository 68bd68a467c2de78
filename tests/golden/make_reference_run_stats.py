"""Writes the fixtures of the reference's own phase-map run (its committed
output/0_STANDALONE/ files, physical_exp_nni_model.py on the 4321 scene on an
A100) into tests/golden/reference_run/:

  stats.json  per model: mean per-Rx coverage in dB (the best code's
              20 log10 |y|, clamped at -70 dB as go.py:944-964 draws it) over
              the 3 x 84 x 75 Rx, the fraction of Rx above -70 dB, and the
              median; from optimization_result.pt's image_list (go.py:1004-1026:
              (0, 0) no code book, (0, 1) code book, (1, 0) code book + MS not
              optimised, (1, 1) code book + MS optimised; each image holds the
              three Rx layers in quadrants (0, 0), (0, 1), (1, 0) and padding in
              (1, 1), go.py:950-956)
  params.npz  the run's optimised phases from ms_and_codebook.pt: ms_theta0
              [32000], ap_noms and ap_ms [30, 36]

Both .pt files are read with torch.load(weights_only=True, map_location="cpu").
Run from the repository root:  python tests/golden/make_reference_run_stats.py
"""
import json
import os

import numpy as np
import torch

REF = "/root/reference/output/0_STANDALONE"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_run")


def layers(img: torch.Tensor) -> torch.Tensor:
    """[168, 150] image -> the three Rx layers [3, 84, 75]"""
    q = img.reshape(2, 84, 2, 75).transpose(1, 2)  # [a, b, 84, 75], layer = 2 a + b
    return q.reshape(4, 84, 75)[:3]


def main():
    os.makedirs(OUT, exist_ok=True)
    d = torch.load(os.path.join(REF, "optimization_result.pt"), weights_only=True, map_location="cpu")
    names = d["names"]
    pos = [(0, 0), (0, 1), (1, 0), (1, 1)]
    stats = {}
    for name, p in zip(names, pos):
        x = layers(d["image_list"][p]).reshape(-1).double()
        stats[name] = {"mean_db": float(x.mean()), "median_db": float(x.median()),
                       "frac_above_floor": float((x > -70).double().mean()), "n_rx": int(x.numel())}
    gain = layers(d["image_list"][(1, 1)]).double() - layers(d["image_list"][(0, 0)]).double()
    stats["gain_codebook_ms_opt_vs_no_codebook"] = {"mean_db": float(gain.mean()), "median_db": float(gain.median())}
    stats["source"] = "output/0_STANDALONE/optimization_result.pt (image_list, names)"
    with open(os.path.join(OUT, "stats.json"), "w") as f:
        json.dump(stats, f, indent=1)
    m = torch.load(os.path.join(REF, "ms_and_codebook.pt"), weights_only=True, map_location="cpu")
    np.savez_compressed(os.path.join(OUT, "params.npz"), **{k: v.detach().numpy() for k, v in m.items()})
    print(json.dumps(stats, indent=1))


if __name__ == "__main__":
    main()
