#!/usr/bin/env python
"""Benchmark of the AutoMS channel and coverage path on one MI355X.

A step is one full channel trace of the workload: forward discovery of
(path, Rx-tile) records, GPU dedup, and the backward evaluation rounds with
neighbour expansion, accumulating complex H[tx, rx].  Inputs (scene handle,
Tx/Rx arrays) are resident before the timed region.

Prints one JSON line:
  value        closest-hit queries per second (Mrays/s), whole job
  rx_channels_per_s   n_tx * n_rx / step seconds, summed over the traces
  roofline     the dominant kernel against its bound, VALU instruction
               issue (wave64 VALU instructions per query from rocprofv3
               PMC x live queries / HIP-event time), with the LDS bytes of
               the staged scene and the HBM traffic beside it
  cpu_baseline the FP64 oracle on a bounded sample of the same workload
  office_full  the office workload at its stated size: 2^24 rays per trace,
               H_d, G_i, F_i of three 48 x 48 metasurfaces and the
               metasurface-to-metasurface traces T_ij, traced concurrently

`python bench.py --pmc-summary DIR` prints the per-kernel table of the
rocprofv3 counter CSV files under DIR (the profiles/ summaries).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
# MFMA flops issued per useful complex-GEMM flop: four split-bf16 real
# products of 3 MFMAs each per complex product
ISSUED_PER_USEFUL = 3.0
# VALU issue: two wave64 VALU instructions per CU per cycle (FP32 and 32-bit
# integer adds issue at 1.86 per CU per cycle in the ams_valu_probe counter
# run, profiles/valu_issue_ceiling.txt; FMA at 0.97, half rate), 256 CUs at
# the 2.4 GHz peak engine clock.  LDS: 128 bytes per clock per CU.
N_CU = 256
PEAK_CLOCK_HZ = 2.4e9
VALU_ISSUE_PER_CU_CYCLE = 2.0
VALU_PEAK_GINST = VALU_ISSUE_PER_CU_CYCLE * N_CU * PEAK_CLOCK_HZ / 1e9
LDS_PEAK_GBS = N_CU * 128 * PEAK_CLOCK_HZ / 1e9

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# in KiB per dispatch, scaled by 2 and 1 (a 1 GiB device copy reads back as
# 0.5 GiB fetched and 1 GiB written), and the L2 hit rate TCC_HIT / (TCC_HIT
# + TCC_MISS); profiles/hbm_traffic.md.
# The evaluation's figures cover each launch pair k_beam + k_evaluate (both
# inside the timed evaluation events): counter sums over the profiled run's
# dispatches of both kernels divided by its evaluation launches.
MEASURED_TRAFFIC = {
    ("room_boxes_ms", "k_discover"): {"bytes": (1.7195e6 * 2048 + 3.0741e6 * 1024) / 6, "l2_hit": 0.891},
    ("room_boxes_ms", "k_evaluate"): {"bytes": (6.3148e5 * 2048 + 1.5575e6 * 1024) / 26, "l2_hit": 0.719},
    ("office", "k_discover"): {"bytes": (2.3763e6 * 2048 + 1.6631e6 * 1024) / 10, "l2_hit": 0.914},
    ("office", "k_evaluate"): {"bytes": (4.4175e7 * 2048 + 9.5146e7 * 1024) / 180, "l2_hit": 0.821},
    ("office_batched", "k_cgemm"): {"bytes": 7.554e5 * 2048 + 1.311e5 * 1024, "l2_hit": 0.782, "mfma_util": 0.589},
}


# VALU wave64 instructions per closest-hit query, from rocprofv3 --pmc
# SQ_INSTS_VALU over a profiled run divided by that run's queries (the
# tracer's counters of the same run); profiles/tracer_pmc_*.txt.
MEASURED_VALU = {
    ("room_boxes_ms", "k_evaluate"): 1.30833e10 / 1517460974,                       # 8.62, k_beam + k_evaluate
    ("room_boxes_ms", "k_discover"): 1.22221e10 / 491237478,                        # 24.88, k_discover<true>
    ("office", "k_evaluate"): 1.5053e12 / 113655676056,                             # 13.24, k_beam + k_evaluate
    ("office", "k_discover"): 6.65919e10 / 1352819735,                              # 49.22, k_discover_wide<false>
}
# VALUBusy (SQ_ACTIVE_INST_VALU per CU over the GUI-active cycles) of the
# same runs: the share of cycles the vector ALUs are busy, whatever each
# instruction's issue cost (FP64, transcendental and 64-bit operations take
# several cycles).
MEASURED_VALU_BUSY = {
    ("room_boxes_ms", "k_evaluate"): 0.844,
    ("room_boxes_ms", "k_discover"): 0.765,
    ("office", "k_evaluate"): 0.994,
    ("office", "k_discover"): 0.664,
}


def pmc_summary(path: str, by_grid: bool = False) -> str:
    """Per-kernel table of the rocprofv3 counter-collection CSV files under
    path: dispatches, mean duration and the mean of every counter per
    dispatch, with the VALU-issue fraction (SQ_INSTS_VALU per cycle per CU
    at the peak clock), VALUBusy, lane utilisation, wait fraction and MFMA
    busy fraction (SQ_VALU_MFMA_BUSY_CYCLES over the GUI-active cycles of
    1024 SIMDs) when those counters are present.  by_grid: one row per
    (kernel, grid size), e.g. to tell two launch shapes of a kernel apart."""
    import csv
    import glob
    import re
    rows = {}
    for fn in sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True)):
        with open(fn) as f:
            for r in csv.DictReader(f):
                name = re.sub(r"\(anonymous namespace\)::", "", r["Kernel_Name"])
                name = re.sub(r"\((?!anonymous).*$", "", name).strip()
                name = re.sub(r"^void ", "", name)
                if by_grid:
                    name += f" [grid {r.get('Grid_Size', '?')}]"
                k = rows.setdefault(name, {"disp": set(), "dur": {}, "c": {}})
                did = r.get("Dispatch_Id") or r.get("Correlation_Id")
                k["disp"].add(did)
                if r.get("Start_Timestamp") and r.get("End_Timestamp"):
                    k["dur"][did] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
                k["c"][r["Counter_Name"]] = k["c"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    out = []
    for name, k in sorted(rows.items(), key=lambda kv: -sum(kv[1]["dur"].values())):
        n = len(k["disp"])
        dur = sum(k["dur"].values()) / max(1, len(k["dur"]))
        c = {a: v / n for a, v in k["c"].items()}
        line = f"{name[:70]:70s} disp {n:5d} avg {dur * 1e3:9.3f} ms"
        for a in sorted(c):
            line += f"  {a} {c[a]:.4g}"
        if "SQ_INSTS_VALU" in c and dur > 0:
            line += (f"  | VALU/CU/cycle {c['SQ_INSTS_VALU'] / (dur * PEAK_CLOCK_HZ * N_CU):.3f} at 2.4 GHz, "
                     f"VALU issue frac {c['SQ_INSTS_VALU'] / (dur * PEAK_CLOCK_HZ * N_CU * VALU_ISSUE_PER_CU_CYCLE):.3f}")
            if c.get("GRBM_GUI_ACTIVE"):
                line += f", {c['SQ_INSTS_VALU'] / N_CU / (c['GRBM_GUI_ACTIVE'] / 8):.3f} per GUI-active cycle"
        if "SQ_ACTIVE_INST_VALU" in c and c.get("GRBM_GUI_ACTIVE"):
            line += f"  VALUBusy {100 * c['SQ_ACTIVE_INST_VALU'] / N_CU / (c['GRBM_GUI_ACTIVE'] / 8):.1f}%"
        if "SQ_THREAD_CYCLES_VALU" in c and c.get("SQ_ACTIVE_INST_VALU"):
            line += f"  lanes {c['SQ_THREAD_CYCLES_VALU'] / (64 * c['SQ_ACTIVE_INST_VALU']):.3f}"
        if "SQ_WAIT_INST_ANY" in c and c.get("SQ_WAVE_CYCLES"):
            line += f"  wait {c['SQ_WAIT_INST_ANY'] / c['SQ_WAVE_CYCLES']:.3f}"
        if "SQ_VALU_MFMA_BUSY_CYCLES" in c and c.get("GRBM_GUI_ACTIVE"):
            line += f"  MFMA busy {c['SQ_VALU_MFMA_BUSY_CYCLES'] / (c['GRBM_GUI_ACTIVE'] / 8 * 4 * N_CU):.3f}"
        out.append(line)
    return "\n".join(out)


def pmc_segments(parent: str) -> str:
    """Tracer kernels of a --profile-pass run, split into its room part and its
    office part (the office starts at the first k_discover_wide dispatch), from
    the four counter runs parent/pmc_valu, pmc_fetch, pmc_write and pmc_l2:
    per kernel the dispatches, mean duration, VALU instructions per dispatch,
    VALU issue fraction, VALUBusy, lane utilisation, wait fraction, HBM MB
    fetched and written per dispatch (FETCH_SIZE x 2, WRITE_SIZE x 1) and the
    L2 hit rate."""
    import collections
    import csv
    import glob
    import re

    def load(path):
        rows = []
        for fn in glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True):
            with open(fn) as f:
                rows += list(csv.DictReader(f))

        def did(r):
            return int(r.get("Dispatch_Id") or r.get("Correlation_Id"))
        first_office = min((did(r) for r in rows if "k_discover_wide" in r["Kernel_Name"]), default=1 << 62)
        agg = collections.defaultdict(lambda: collections.defaultdict(float))
        dur = collections.defaultdict(dict)
        for r in rows:
            name = re.sub(r"\(anonymous namespace\)::", "", r["Kernel_Name"])
            name = re.sub(r"^void ", "", re.sub(r"\((?!anonymous).*$", "", name).strip())
            if not name.startswith(("k_beam", "k_evaluate", "k_discover")):
                continue
            key = ("office" if did(r) >= first_office else "room", name)
            agg[key][r["Counter_Name"]] += float(r["Counter_Value"])
            dur[key][did(r)] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
        return agg, dur

    runs = {p: load(os.path.join(parent, "pmc_" + p)) for p in ("valu", "fetch", "write", "l2")}
    agg, dur = runs["valu"]
    out = []
    for key in sorted(agg):
        c, n, t = agg[key], len(dur[key]), sum(dur[key].values())
        fe = runs["fetch"][0][key]["FETCH_SIZE"] * 2048 / n / 1e6
        wr = runs["write"][0][key]["WRITE_SIZE"] * 1024 / n / 1e6
        h = runs["l2"][0][key]
        out.append(f"{key[0]:6s} {key[1]:38s} disp {n:4d} avg {t / n * 1e3:8.3f} ms  "
                   f"VALU/disp {c['SQ_INSTS_VALU'] / n:.4g}  "
                   f"issue {c['SQ_INSTS_VALU'] / (t * PEAK_CLOCK_HZ * N_CU * VALU_ISSUE_PER_CU_CYCLE):.3f}  "
                   f"VALUBusy {c['SQ_ACTIVE_INST_VALU'] / N_CU / (c['GRBM_GUI_ACTIVE'] / 8):.3f}  "
                   f"lanes {c['SQ_THREAD_CYCLES_VALU'] / (64 * c['SQ_ACTIVE_INST_VALU']):.3f}  "
                   f"wait {c['SQ_WAIT_INST_ANY'] / c['SQ_WAVE_CYCLES']:.3f}  fetch {fe:.1f} MB  write {wr:.1f} MB  "
                   f"L2 hit {h['TCC_HIT'] / (h['TCC_HIT'] + h['TCC_MISS']):.3f}")
    return "\n".join(out)


def kt_segments(parent: str, warmup: int, steps: int, traces: int = 3) -> str:
    """Tracer kernels of a rocprofv3 --kernel-trace run of this script, split
    by the first workload's discovery launches (`traces` per step) into the
    timed steps (after `warmup` warm-up steps) and the isolated step that
    follows them: per kernel the dispatches, mean duration and time per
    step, and the evaluation (k_beam + k_evaluate) per k_evaluate dispatch,
    the unit of the bench line's live evaluation figure."""
    import collections
    import csv
    import glob
    import re

    recs = []
    for fn in glob.glob(os.path.join(parent, "**", "*kernel_trace.csv"), recursive=True):
        with open(fn) as f:
            for r in csv.DictReader(f):
                name = re.sub(r"\(anonymous namespace\)::", "", r["Kernel_Name"])
                name = re.sub(r"^void ", "", re.sub(r"\((?!anonymous).*$", "", name).strip())
                if name.startswith(("k_beam", "k_evaluate", "k_discover")):
                    recs.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), name))
    recs.sort()
    disc_name = next(n for _, _, n in recs if n.startswith("k_discover"))
    disc = [t for t, _, n in recs if n == disc_name]
    k0, k1, k2 = traces * warmup, traces * (warmup + steps), traces * (warmup + steps + 1)
    if len(disc) < k2:
        raise SystemExit(f"{len(disc)} {disc_name} dispatches: fewer than {k2} (warm-up, timed and isolated steps)")
    segs = {"timed": (disc[k0], disc[k1], steps), "isolated": (disc[k1], disc[k2] if len(disc) > k2 else 1 << 62, 1)}
    out = []
    for seg, (lo, hi, n_steps) in segs.items():
        agg = collections.defaultdict(list)
        for t0, t1, n in recs:
            if lo <= t0 < hi:
                agg[n].append((t1 - t0) * 1e-6)
        for n in sorted(agg):
            d = agg[n]
            out.append(f"{seg:8s} {n:40s} dispatches {len(d):4d}  avg {sum(d) / len(d):8.3f} ms  "
                       f"per step {sum(d) / n_steps:8.3f} ms")
        ev = [x for n in agg if n.startswith("k_evaluate") for x in agg[n]]
        bm = [x for n in agg if n.startswith("k_beam") for x in agg[n]]
        if ev:
            out.append(f"{seg:8s} {'evaluation (k_beam + k_evaluate)':40s} dispatches {len(ev):4d}  "
                       f"avg {(sum(ev) + sum(bm)) / len(ev):8.3f} ms  per step {(sum(ev) + sum(bm)) / n_steps:8.3f} ms")
    return "\n".join(out)


def valu_roofline(workload: str, kernel: str, queries: float, launches: int, kernel_ms: float,
                  alg_bytes: float, lds_staged: bool) -> dict:
    """The roofline object of a tracer kernel: VALU instruction issue (the
    bound: about 30 instructions per node visit and lanes 96% active), with
    the LDS bytes of the staged scene and the measured HBM traffic beside
    it.  Live: launches, queries and HIP-event time of the timed steps."""
    per_launch_s = kernel_ms * 1e-3 / max(1, launches)
    ipq = MEASURED_VALU.get((workload, kernel))
    tr = MEASURED_TRAFFIC.get((workload, kernel), {})
    out = {"kernel": kernel, "bound": "valu", "unit": "Ginst/s", "peak": VALU_PEAK_GINST,
           "avg_launch_ms": per_launch_s * 1e3, "queries_per_launch": queries / max(1, launches),
           "valu_inst_per_query": ipq, "valu_busy_pmc": MEASURED_VALU_BUSY.get((workload, kernel)),
           "achieved": None, "frac": None, "traffic": tr.get("bytes"), "l2_hit_rate": tr.get("l2_hit")}
    if kernel == "k_evaluate":
        out["launch"] = ("k_beam (candidate lists, Tx arrays of 256 elements and more) + k_evaluate, timed together "
                         "by the evaluation events")
    if ipq is not None:
        out["achieved"] = ipq * queries / max(1, launches) / per_launch_s / 1e9
        out["frac"] = out["achieved"] / VALU_PEAK_GINST
    if tr.get("bytes"):
        out["hbm"] = {"gb_per_s": tr["bytes"] / per_launch_s / 1e9, "frac": tr["bytes"] / per_launch_s / 1e9 /
                      HBM_PEAK_GBS}
    abps = alg_bytes / max(1, launches) / per_launch_s / 1e9
    if lds_staged:
        out["lds"] = {"algorithmic_gb_per_s": abps, "peak": LDS_PEAK_GBS, "frac": abps / LDS_PEAK_GBS,
                      "what": "32 B per node visit and 48 B per triangle test, read from the LDS-staged scene"}
    else:
        out["scene_fetch"] = {"algorithmic_gb_per_s": abps,
                              "what": "32 B per node visit and 48 B per triangle test, served by L2 and MALL"}
    return out


def step_valu(workload: str, tot: dict, seconds: float) -> dict:
    """VALU issue of the whole timed region: the discovery and evaluation
    kernels' instructions (measured per query) over its wall time, i.e. the
    GPU's VALU use with the concurrent traces sharing it."""
    ins = (MEASURED_VALU[(workload, "k_discover")] * tot["fwd_queries"] +
           MEASURED_VALU[(workload, "k_evaluate")] * tot["bwd_queries"])
    return {"achieved": ins / seconds / 1e9, "frac": ins / seconds / 1e9 / VALU_PEAK_GINST,
            "what": "VALU instructions of every discovery and evaluation launch of the timed steps "
                    "(per-query counts from the profiled run) over the steps' wall time"}


def kernel_bytes(stats: dict) -> dict:
    """Algorithmic bytes per kernel: 32 B per BVH node visit, 48 B per
    triangle test, 16 B per sphere test, 64 B of ray / hit I/O per query."""
    fwd = 32 * stats["fwd_nodes"] + 48 * stats["fwd_tris"] + 16 * stats["fwd_spheres"] + 64 * stats["fwd_queries"]
    bwd = 32 * stats["bwd_nodes"] + 48 * stats["bwd_tris"] + 64 * stats["bwd_queries"]
    return {"k_discover": fwd, "k_evaluate": bwd}


def workload_traces(name: str):
    from automs_amd import scenes
    if name == "room_boxes":
        return {"H": scenes.room_boxes()}
    if name == "room_boxes_ms":
        return scenes.room_boxes_ms()
    if name == "office":  # 100,046 triangles, 5 reflections: H_d, G0, F0 of one metasurface, 2^24 rays each
        tr = scenes.office()
        return {k: tr[k] for k in ("H_d", "G0", "F0")}
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(name: str, sample_fib: int, sample_ms_fib: int) -> dict:
    """FP64 oracle (OpenMP) on the same scene with fewer rays per trace."""
    from automs_amd import scenes
    from oracle import oracle as O
    if name == "office":
        w = scenes.office(fib_n=1 << 10, ms_fib_n=1 << 12)["H_d"]
        t0 = time.perf_counter()
        _, st, _ = O.trace_workload(w)
        dt = time.perf_counter() - t0
        q = st["fwd_queries"] + st["bwd_queries"]
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        return {"value": q / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
                "sample": f"office H_d ({w.scene.n_tri} triangles, 64 x 65,536) with {w.fib_n} discovery rays "
                          f"({q} closest-hit queries in {dt:.1f} s; FP64 oracle, OpenMP)"}
    if name == "room_boxes":
        traces = {"H": scenes.room_boxes(fib_n=sample_fib)}
    else:
        traces = scenes.room_boxes_ms(fib_n=sample_fib, ms_fib_n=sample_ms_fib, ms_size=(16, 16))
    q = 0
    t0 = time.perf_counter()
    for w in traces.values():
        _, st, _ = O.trace_workload(w)
        q += st["fwd_queries"] + st["bwd_queries"]
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    what = f"{name}: every trace with {sample_fib} Fibonacci rays per AP element"
    if name != "room_boxes":
        what += f", a 16 x 16 metasurface and {sample_ms_fib} rays in total from it"
    out = {"value": q / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
           "sample": f"{what} ({q} closest-hit queries in {dt:.1f} s; FP64 oracle, OpenMP)"}
    return out


def cpu_phase_map(channels: dict, max_phase_rad: float, target_db: float, tol_db: float = 0.1,
                  max_steps: int = 120, code_num: int = 30, seed: int = 0) -> dict:
    """Host baseline of the phase-map optimisation: the torch-CPU restatement
    of go.py's Model (oracle.contract_ref: H_d + G diag(u) F, then the AP
    codebook) with torch.optim.Adam (lr 0.3), on all host threads, in FP32
    (go.py's precision) and FP64, from the same seeded start as the GPU run.
    Runs until the mean best-code power is within tol_db of target_db (the
    plain FP32 GPU optimum) or max_steps; the same step counting as the GPU
    measurement (a step's coverage is that of its forward)."""
    import torch

    from automs_amd.phasemap import N0_90DBM, ap_amplitude, start_phases
    from oracle import contract_ref as R
    Hd = channels["src_to_dst"].cpu()
    G = channels["src_to_ms{0}"].cpu()
    F = channels["ms{0}_to_dst"].cpu()
    M, K = G.shape
    th_ms0, th_ap0 = start_phases(K, M, code_num, seed)
    amp = ap_amplitude(code_num, M)
    out = {"threads": torch.get_num_threads(), "kind": "port",
           "what": "torch-CPU restatement of go.py's Model + torch.optim.Adam (oracle.contract_ref)"}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        cdt = torch.complex64 if dt == torch.float32 else torch.complex128
        Hd_, G_, F_ = Hd.to(cdt), G.to(cdt), F.to(cdt)
        ta = th_ap0.to(dt).clone().requires_grad_(True)
        tm = th_ms0.reshape(-1).to(dt).clone().requires_grad_(True)
        opt = torch.optim.Adam([ta, tm], lr=0.3)
        hit = None
        t0 = time.perf_counter()
        for it in range(max_steps):
            Y = R.received(amp, ta, Hd_, [G_], [F_], [tm], [max_phase_rad], dtype=dt)
            loss = R.loss_logsnr(Y, 1, N0_90DBM)
            if float(R.coverage_db(Y.detach())[0]) >= target_db - tol_db:
                hit = it + 1
                break
            opt.zero_grad()
            loss.backward()
            opt.step()
        dt_s = time.perf_counter() - t0
        n = hit if hit is not None else max_steps
        out[name] = {"steps_to_0p1dB": hit, "seconds_to_0p1dB": dt_s if hit is not None else None,
                     "seconds_per_step": dt_s / n, "steps_run": n}
    return out


def contraction_office(iters: int = 20) -> dict:
    """Config-3 shape on traced office channels: 4 phase maps x 64 AP beams
    (M = 256) against [F0 ; H_d] (K = 2304 + 64, N = 65,536).  Times the
    forward split-bf16 MFMA GEMM alone (HIP events on the current stream) and
    a full optimiser step (forward, objective, hand-written backward, Adam)."""
    import torch

    from automs_amd import scenes
    from automs_amd.phasemap import Contract, FusedAdam, Metasurface, batched_model, cgemm_frag, frag_pack
    from automs_amd.tracer import TraceScene, trace, trace_many
    tr = scenes.office(fib_n=1 << 20, ms_fib_n=1 << 20)
    ch = {}
    traced = {}
    for key, name in (("src_to_dst", "H_d"), ("src_to_ms{0}", "G0"), ("ms{0}_to_dst", "F0")):
        w = tr[name]
        res, st, _ = trace(TraceScene(w.scene), w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                           w.gridsize)
        ch[key] = torch.complex(res[..., 0], res[..., 1])
        q = st["fwd_queries"] + st["bwd_queries"]
        traced[name] = {"n_tx": w.tx.n, "n_rx": w.rx.n, "rays": st["rays"], "queries": q, "paths": st["paths"],
                        "ms": st["t_total_ms"], "mrays_per_s": q / (st["t_total_ms"] * 1e-3) / 1e6,
                        "rx_channels_per_s": w.tx.n * w.rx.n / (st["t_total_ms"] * 1e-3)}
    model = batched_model(ch, Metasurface(None, (48, 48), None, None, [], max_phase=360.0))
    op = model.op
    with torch.no_grad():
        A = model.stacked_operand().contiguous()
    M, Kc, N = A.shape[0], op.Kc, op.N
    fa = frag_pack(A, M, Kc, Kc, 1)
    for _ in range(3):
        cgemm_frag(fa, op.frag_fwd, M, N, Kc, conj_b=False)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        cgemm_frag(fa, op.frag_fwd, M, N, Kc, conj_b=False)
    e1.record()
    torch.cuda.synchronize()
    t_fwd = e0.elapsed_time(e1) / iters * 1e-3
    e0.record()
    for _ in range(iters):
        cgemm_frag(fa, op.frag_fwd, M, N, Kc, conj_b=False, three_mult=True)
    e1.record()
    torch.cuda.synchronize()
    t_fwd_3m = e0.elapsed_time(e1) / iters * 1e-3
    useful = 8.0 * M * Kc * N
    # the backward product gA = gY . conj([F0 ; H_d])^T alone, split over Rx
    g = torch.Generator(device=A.device).manual_seed(3)
    gY = torch.complex(torch.randn(M, N, generator=g, device=A.device), torch.randn(M, N, generator=g, device=A.device))
    fg = frag_pack(gY, M, N, N, 1)
    for _ in range(3):
        cgemm_frag(fg, op.frag_bwd, M, Kc, N, conj_b=True, splits=op.bwd_splits)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(iters):
        cgemm_frag(fg, op.frag_bwd, M, Kc, N, conj_b=True, splits=op.bwd_splits)
    e1.record()
    torch.cuda.synchronize()
    t_bwd = e0.elapsed_time(e1) / iters * 1e-3
    e0.record()
    for _ in range(iters):
        cgemm_frag(fg, op.frag_bwd, M, Kc, N, conj_b=True, splits=op.bwd_splits, split_minor=True)
    e1.record()
    torch.cuda.synchronize()
    t_bwd_minor = e0.elapsed_time(e1) / iters * 1e-3
    e0.record()
    for _ in range(iters):
        cgemm_frag(fg, op.frag_bwd, M, Kc, N, conj_b=True, splits=op.bwd_splits, three_mult=True)
    e1.record()
    torch.cuda.synchronize()
    t_bwd_3m = e0.elapsed_time(e1) / iters * 1e-3
    del gY, fg
    opt = FusedAdam([p for p in model.parameters() if p.requires_grad], lr=0.3)
    for _ in range(2):
        model().backward()
        opt.step()
        opt.zero_grad()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(iters):
        model().backward()
        opt.step()
        opt.zero_grad()
    e1.record()
    torch.cuda.synchronize()
    t_step = e0.elapsed_time(e1) / iters * 1e-3
    tr_meas = MEASURED_TRAFFIC[("office_batched", "k_cgemm")]
    # parity against complex128 on a column sample, and the host baseline:
    # the same contraction in complex64 with torch on all host cores
    with torch.no_grad():
        Y = cgemm_frag(fa, op.frag_fwd, M, N, Kc, conj_b=False)
        cols = torch.arange(0, N, 32, device=Y.device)
        Fcat = torch.cat([ch["ms{0}_to_dst"], ch["src_to_dst"]], 0)
        Fs = Fcat[:, cols].cpu()
        A_h = A.cpu()
        ref = A_h.to(torch.complex128) @ Fs.to(torch.complex128)
        err = (Y[:, cols].cpu().to(torch.complex128) - ref).abs().pow(2).sum(0).sqrt() / \
            ref.abs().pow(2).sum(0).sqrt().clamp_min(1e-300)
        # the same product with the GPU's FP32 library GEMM (complex64 matmul)
        Ag = A.to(torch.complex64).contiguous()
        Yl = Ag @ Fcat
        torch.cuda.synchronize()
        e0.record()
        for _ in range(iters):
            Yl = Ag @ Fcat
        e1.record()
        torch.cuda.synchronize()
        t_lib = e0.elapsed_time(e1) / iters * 1e-3
        err_lib = (Yl[:, cols].cpu().to(torch.complex128) - ref).abs().pow(2).sum(0).sqrt() / \
            ref.abs().pow(2).sum(0).sqrt().clamp_min(1e-300)
        del Yl
        # the issued work of the split-bf16 kernel as one real bf16 GEMM on
        # the library (24 M K N flops: rows [re; im] of A's six split planes
        # against the matching planes of B), in the library's fastest layout
        # here (the long dimension as rows): the practical MFMA ceiling for
        # that work at the clock the chip holds under it
        ga = torch.randn(N, 6 * Kc, device=Y.device, dtype=torch.bfloat16)
        gb = torch.randn(6 * Kc, 2 * M, device=Y.device, dtype=torch.bfloat16)
        for _ in range(3):
            gc = ga @ gb
        torch.cuda.synchronize()
        e0.record()
        for _ in range(iters):
            gc = ga @ gb
        e1.record()
        torch.cuda.synchronize()
        t_ceil = e0.elapsed_time(e1) / iters * 1e-3
        del ga, gb, gc
        Fcpu = Fcat[:, : min(N, 8192)].cpu().contiguous()
        A_h @ Fcpu
        t0 = time.perf_counter()
        A_h @ Fcpu
        t_cpu = (time.perf_counter() - t0) * N / Fcpu.shape[1]
    parity = {"rx_sampled": int(cols.numel()), "per_rx_rel_err_median": float(err.median()),
              "per_rx_rel_err_max": float(err.max()),
              "fro_rel_err": float((Y[:, cols].cpu().to(torch.complex128) - ref).norm() / ref.norm()),
              "oracle": "complex128 product of the same operands on the host"}
    host = {"seconds_per_contraction": t_cpu, "threads": torch.get_num_threads(),
            "kind": "torch complex64 matmul on the host (FP32), timed on 8192 Rx columns and scaled to N",
            "gpu_speedup": t_cpu / t_fwd}
    lib = {"ms": t_lib * 1e3, "per_rx_rel_err_max": float(err_lib.max()), "speedup_of_mfma_path": t_lib / t_fwd,
           "kind": "torch complex64 matmul on the GPU (the FP32 library GEMM), same operands"}
    return {"shape": {"M": M, "K": Kc, "N": N}, "fwd_gemm_ms": t_fwd * 1e3,
            "bwd_gemm_ms": t_bwd * 1e3, "bwd_splits": op.bwd_splits, "bwd_useful_tflops": useful / t_bwd / 1e12,
            "bwd_gemm_ms_split_minor": t_bwd_minor * 1e3,
            "three_product_form": {"fwd_gemm_ms": t_fwd_3m * 1e3, "bwd_gemm_ms": t_bwd_3m * 1e3,
                                   "what": "the workgroup kernel with 9 MFMAs per complex tile pair instead of 12 "
                                           "(six planes per block staged in LDS)"},
            "rx_channels_per_s": M * N / t_fwd, "parity_vs_fp64": parity, "host_baseline": host,
            "gpu_fp32_library_gemm": lib,
            "library_bf16_same_issued_work": {
                "ms": t_ceil * 1e3, "tflops": 24.0 * M * Kc * N / t_ceil / 1e12, "fwd_over_library": t_fwd / t_ceil,
                "kind": f"torch bf16 matmul (hipBLASLt) [{N} x {6 * Kc}] @ [{6 * Kc} x {2 * M}], random operands: the "
                        "forward's issued flops as one real GEMM"},
            "roofline": {"bound": "mfma", "achieved": ISSUED_PER_USEFUL * useful / t_fwd / 1e12, "peak": MFMA_BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": ISSUED_PER_USEFUL * useful / t_fwd / 2.5e15, "traffic": tr_meas["bytes"],
                         "mfma_busy_pmc": tr_meas["mfma_util"]},
            "useful_tflops": useful / t_fwd / 1e12, "issued_tflops": ISSUED_PER_USEFUL * useful / t_fwd / 1e12,
            "mfma_bf16_peak_tflops": 2500.0, "mfma_frac": ISSUED_PER_USEFUL * useful / t_fwd / 2.5e15,
            "f_operand_gb_per_s": 16.0 * Kc * N / t_fwd / 1e9,
            "optimizer_step_ms": t_step * 1e3, "maps": 4, "beams": 64,
            "office_traces": {"scene_tris": tr["H_d"].scene.n_tri, "max_reflection": tr["H_d"].max_reflection,
                              "traces": traced}}


def reference_run() -> dict:
    """The reference's own phase-map run (physical_exp_nni_model.py on the
    4321 scene: 6 x 6 AP, 3 x 84 x 75 Rx, one 160 x 200 metasurface, 4
    reflections, the reference's rays; four traces, then three models of
    200 Adam steps), end to end, against its published time (about 140 s on
    an A100 including its plots, README.md:140) and, loosely, its mean
    coverage figures (tests/golden/reference_run, written from its
    output/0_STANDALONE files)."""
    import numpy as np

    from automs_amd import experiment as E
    folder = str(ROOT / "tests" / "golden" / "scene_4321")
    r = E.run_trial(folder)
    ref = json.load(open(ROOT / "tests" / "golden" / "reference_run" / "stats.json"))
    p = np.load(ROOT / "tests" / "golden" / "reference_run" / "params.npz")
    ch = r["_channels"]
    _, _, _, ms = E.trial_geometry(folder)
    with_ref = {E.MODEL_NAMES[1]: E.coverage_with_phases(ch["noms"], [], p["ap_noms"]),
                E.MODEL_NAMES[3]: E.coverage_with_phases(ch["ms"], [ms], p["ap_ms"], p["ms_theta0"].reshape(1, -1))}
    return {
        "workload": "physical_exp_nni_model.py: 4321 scene (2689 triangles), 6 x 6 AP, 3 x 84 x 75 Rx, a 160 x 200 "
                    "metasurface on wall 1 at ms_j 0.1194, 4 reflections, 100 x 100 reference rays per trace; "
                    "models: AP codebook, + MS fixed, + MS optimised, 200 Adam steps each (capacity loss, w 0.5)",
        "seconds_total": r["seconds_total"], "seconds_traces": r["seconds_traces"],
        "seconds_optimise": r["seconds_optimise"], "published_seconds": 140.0,
        "published": "about 140 s on an NVIDIA A100, plots included (README.md:140)",
        "speedup_vs_published": 140.0 / r["seconds_total"],
        "mean_coverage_db": {k: v["mean_db"] for k, v in r["models"].items()},
        "reference_mean_coverage_db": {k: ref[k]["mean_db"] for k in E.MODEL_NAMES},
        "frac_rx_above_minus_70_db": {k: v["frac_above_floor"] for k, v in r["models"].items()},
        "reference_frac_rx_above_minus_70_db": {k: ref[k]["frac_above_floor"] for k in E.MODEL_NAMES},
        "reference_phases_on_these_channels_mean_db": {k: v["mean_db"] for k, v in with_ref.items()},
        "final_result": r["final_result"],
        "shapes": r["shapes"],
    }


def surrogate_training(n: int = 2048, epochs: int = 2, batch: int = 64) -> dict:
    """Training throughput of the placement surrogate at its stated shape
    (9 input channels, four 3 x 3 convolutions of 32 channels, 64 x 64 maps,
    batch 64), on synthetic inputs and labels from the same start: the bf16
    MFMA kernels against the FP32 vector-ALU kernels.  Work per sample and
    step: the forward, the input gradients of layers 1-3 and every weight
    gradient, 2 flops per multiply-add."""
    import torch

    from automs_amd import surrogate as S
    g = torch.Generator(device="cuda").manual_seed(0)
    X = torch.randn(n, 9, 64, 64, generator=g, device="cuda")
    Y = torch.randn(n, 1, 64, 64, generator=g, device="cuda")
    ch = (9, 32, 32, 32, 1)
    macs = [ch[i] * 9 * ch[i + 1] * 64 * 64 for i in range(4)]
    flop = 2.0 * (sum(macs) + sum(macs[1:]) + sum(macs))
    out = {"samples": n, "epochs": epochs, "batch": batch, "flop_per_sample_step": flop}
    for mfma in (True, False):
        net = S.SurrogateNet(seed=0, mfma=mfma)
        S.train(net, X[:2 * batch], Y[:2 * batch], epochs=1, batch=batch)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hist = S.train(net, X, Y, epochs=epochs, batch=batch)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ss = epochs * (n // batch) * batch
        out["mfma" if mfma else "fp32"] = {"seconds": dt, "sample_steps_per_s": ss / dt,
                                           "tflops": ss * flop / dt / 1e12, "final_mse": hist[-1]}
    out["speedup_mfma_over_fp32"] = out["fp32"]["seconds"] / out["mfma"]["seconds"]
    return out


def placement_search(sa_steps: int) -> dict:
    """The placement search (hyper-config tuner of physical_exp_nni_main.py
    with its trial objective): plain simulated annealing with every candidate
    judged by a traced and optimised trial, against annealing steered by the
    conv-net surrogate with the tracer re-scoring every accepted candidate,
    on a room scene outside the training set (seed 1000).  The network is
    the committed one (automs_amd/data/surrogate_room_trial.*: its dataset
    size, training and validation are in the .json next to it)."""
    from automs_amd import surrogate as S
    data = ROOT / "automs_amd" / "data"
    net = S.SurrogateNet()
    net.load(str(data / "surrogate_room_trial.npy"))
    r = S.search_experiment(net, sa_steps=sa_steps, search_scene=1000, rank_n=64,
                            log=lambda m: print(m, file=sys.stderr, flush=True))
    meta = json.load(open(data / "surrogate_room_trial.json"))
    r["network"] = {k: meta[k] for k in ("dataset", "training", "validation", "command")}
    return r


def office_phase_map(outs: dict, steps: int = 200) -> dict:
    """The phase-map optimisation of config 2's three metasurfaces on the
    traced office channels (go.py's Model with max_ms_reflex 2, its default):
    30 codes of the 8 x 8 AP, chains over the three 48 x 48 surfaces through
    the ms-to-ms channels, mean log-SNR over the 65,536 Rx, Adam lr 0.3, the
    step replayed from a HIP graph; against the codebook optimised alone."""
    import torch

    from automs_amd import experiment, scenes
    from automs_amd.phasemap import Metasurface, PhaseMapModel, optimize_model
    ch = {scenes.OFFICE_CHANNEL_KEYS[k]: torch.complex(v[..., 0], v[..., 1]) for k, v in outs.items()}
    ms = [Metasurface(None, (48, 48), None, None, [], max_phase=360.0) for _ in range(3)]
    res = {}
    for name, chans, surfaces in (("codebook_alone", {"src_to_dst": ch["src_to_dst"]}, []),
                                  ("three_surfaces", ch, ms)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        model = PhaseMapModel(name, chans, surfaces, code_num=30, objective="logsnr", max_ms_reflex=2, seed=0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        st0 = {f: experiment.coverage_stats(model, f) for f in (-70.0, -100.0)}
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        optimize_model(model, lr=0.3, step=steps, graph=True)
        e1.record()
        torch.cuda.synchronize()
        with torch.no_grad():
            model()
        st1 = {f: experiment.coverage_stats(model, f) for f in (-70.0, -100.0)}
        res[name] = {"setup_seconds": t1 - t0, "seconds": e0.elapsed_time(e1) * 1e-3,
                     "ms_per_step": e0.elapsed_time(e1) / steps, "loss_end": float(model.loss.item()),
                     "best_loss": float(min(model.loss_list)),
                     "coverage_start": {f"floor_{-int(f)}_db": {"mean_db": v["mean_db"], "frac_above": v["frac_above_floor"]}
                                        for f, v in st0.items()},
                     "coverage_end": {f"floor_{-int(f)}_db": {"mean_db": v["mean_db"], "frac_above": v["frac_above_floor"]}
                                      for f, v in st1.items()}}
        if surfaces:
            res[name]["operand_gb"] = model.op.nbytes() / 1e9
            res[name]["chain_transfer_gb"] = model.chain.frag.numel() / 1e9
        del model
        torch.cuda.empty_cache()
    res["coverage_gain_db"] = {k: res["three_surfaces"]["coverage_end"][k]["mean_db"]
                               - res["codebook_alone"]["coverage_end"][k]["mean_db"]
                               for k in res["three_surfaces"]["coverage_end"]}
    res["coverage"] = ("per Rx the best code's 20 log10 |y|, clamped at the floor (go.py:944-964's rule with its "
                       "-70 dB floor, and -100 dB), averaged over the 65,536 Rx")
    res["workload"] = ("config 2's traced office channels: 8 x 8 AP, 30 codes, three 48 x 48 metasurfaces, chains "
                       "of up to 2 surfaces, 65,536 Rx; K = 3 x 2304 + 64, the split-bf16 operand in both layouts")
    res["steps"] = steps
    return res


def office_full(steps: int, warmup: int, cpu: bool, profile: bool = False) -> dict:
    """The office workload at its stated size (100,046 triangles, 5
    reflections, 2^24 rays per trace): H_d [64 x 65,536], G_i [64 x 2304],
    F_i [2304 x 65,536] of three 48 x 48 metasurfaces and T_ij [2304 x 2304],
    the ten traces of build_channels (go.py:1135-1247), traced concurrently
    by tracer.trace_many.  A step is all ten traces; then one untimed step
    with the traces in turn gives each kernel's launch durations alone on the
    GPU."""
    import torch

    from automs_amd import scenes
    from automs_amd.tracer import TraceScene, trace, trace_many
    dev = torch.device("cuda:0")
    tr = scenes.office()
    handles = {k: TraceScene(w.scene) for k, w in tr.items()}
    outs = {k: torch.zeros(w.tx.n, w.rx.n, 2, dtype=torch.float32, device=dev) for k, w in tr.items()}

    def job(k):
        w = tr[k]
        return dict(ts=handles[k], tx=w.tx, rx=w.rx, wave_len=w.wave_len, max_reflection=w.max_reflection,
                    fib_n=w.fib_n, per_element=w.per_element, gridsize=w.gridsize, result=outs[k])

    def step():
        for k in tr:
            outs[k].zero_()
        return {k: r[1] for k, r in zip(tr, trace_many([job(k) for k in tr]))}

    for i in range(warmup):
        step()
        print(f"office warmup {i + 1}/{warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per = {k: {} for k in tr}
    for i in range(steps):
        sts = step()
        for k, st in sts.items():
            for f, v in st.items():
                per[k][f] = per[k].get(f, 0) + v
        print(f"office step {i + 1}/{steps}: {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tot = {}
    for k in tr:
        for f, v in per[k].items():
            tot[f] = tot.get(f, 0) + v
    queries = tot["fwd_queries"] + tot["bwd_queries"]
    iso = {}
    if not profile:
        for k in tr:
            outs[k].zero_()
            _, st, _ = trace(**job(k))
            for f, v in st.items():
                iso[f] = iso.get(f, 0) + v
        torch.cuda.synchronize()
    ref = iso or tot
    ikt = {"k_discover": ref["t_discover_ms"], "k_evaluate": ref["t_evaluate_ms"]}
    dom = max(ikt, key=ikt.get)
    q_of = {"k_discover": "fwd_queries", "k_evaluate": "bwd_queries"}
    l_of = {"k_discover": "n_discover_launches", "k_evaluate": "n_evaluate_launches"}
    t_of = {"k_discover": "t_discover_ms", "k_evaluate": "t_evaluate_ms"}
    kb = kernel_bytes(tot)
    roof = valu_roofline("office", dom, tot[q_of[dom]], tot[l_of[dom]], tot[t_of[dom]], kb[dom], False)
    if iso:
        ikb = kernel_bytes(iso)
        r_iso = valu_roofline("office", dom, iso[q_of[dom]], iso[l_of[dom]], iso[t_of[dom]], ikb[dom], False)
        roof["isolated"] = {k: r_iso[k] for k in ("avg_launch_ms", "achieved", "frac")}
        roof["isolated"]["step_seconds"] = (iso["t_total_ms"]) * 1e-3
    roof["step"] = step_valu("office", tot, dt)
    n_channels = sum(w.tx.n * w.rx.n for w in tr.values())
    out = {
        "workload": "office, 100,046 triangles, <= 5 reflections, 8 x 8 AP, 3 metasurfaces of 48 x 48, "
                    "256 x 256 Rx, 2^24 rays per trace; traces H_d, G0-2, F0-2, T01, T02, T12 concurrently",
        "steps": steps, "warmup": warmup, "seconds_per_step": dt / steps,
        "mrays_per_s": queries / dt / 1e6, "rx_channels_per_s": n_channels * steps / dt,
        "primary_rays_per_s": tot["rays"] / dt,
        "traces": {k: {"n_tx": tr[k].tx.n, "n_rx": tr[k].rx.n, "rays": per[k]["rays"] // steps,
                       "paths": per[k]["paths"] // steps, "successes": per[k]["successes"] // steps,
                       "queries": (per[k]["fwd_queries"] + per[k]["bwd_queries"]) // steps,
                       "wall_ms": per[k]["t_total_ms"] / steps} for k in tr},
        "roofline": roof,
        "totals": {f: tot[f] for f in ("fwd_queries", "bwd_queries", "fwd_nodes", "bwd_nodes",
                                       "n_discover_launches", "n_evaluate_launches")},
    }
    if not profile:
        for k in handles:
            handles[k].close()
        out["phase_map"] = office_phase_map(outs)
        print(f"office phase map: {out['phase_map']}", file=sys.stderr, flush=True)
    if cpu:
        from oracle import oracle as O
        w = scenes.office(fib_n=1 << 10, ms_fib_n=1 << 12)["H_d"]
        t1 = time.perf_counter()
        _, st, _ = O.trace_workload(w)
        dto = time.perf_counter() - t1
        qo = st["fwd_queries"] + st["bwd_queries"]
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": qo / dto / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
                               "seconds": dto, "rx_channels_per_s": w.tx.n * w.rx.n / dto,
                               "sample": f"office H_d ({w.scene.n_tri} triangles, 64 x 65,536) with {w.fib_n} "
                                         f"discovery rays ({qo} closest-hit queries in {dto:.1f} s; FP64 oracle, "
                                         "OpenMP); the neighbour expansion then traces every reachable tile"}
        out["gpu_over_cpu_mrays"] = out["mrays_per_s"] / out["cpu_baseline"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 10 timed steps (about 0.35 s of tracing): the three traces of a step run
    # concurrently, and their interleaving varies from step to step
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="room_boxes_ms", choices=["room_boxes", "room_boxes_ms", "office"])
    ap.add_argument("--opt-steps", type=int, default=500)
    ap.add_argument("--no-opt", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="run the phase-map steps eagerly, not as a HIP graph")
    ap.add_argument("--sequential", action="store_true",
                    help="trace H_d, G and F one after the other instead of concurrently (tracer.trace_many)")
    ap.add_argument("--no-contract", action="store_true")
    ap.add_argument("--trace-flags", type=int, default=0,
                    help="ams_trace_opts.flags for every trace of the timed workload (A/B of code paths)")
    ap.add_argument("--profile-pass", action="store_true",
                    help="short run for rocprof counter passes: 1 traced step, 5 contraction GEMMs and a "
                         "1 GiB device copy that calibrates FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-office", action="store_true", help="skip the office workload at its stated size")
    ap.add_argument("--no-reference-run", action="store_true", help="skip the reference's own phase-map run")
    ap.add_argument("--no-placement-search", action="store_true", help="skip the surrogate placement search")
    ap.add_argument("--only-contract", action="store_true", help="only the config-3 contraction timings")
    ap.add_argument("--only-surrogate", action="store_true", help="only the surrogate training throughput")
    ap.add_argument("--sa-steps", type=int, default=2000, help="annealing steps of each placement search")
    ap.add_argument("--office-steps", type=int, default=2)
    ap.add_argument("--office-warmup", type=int, default=1)
    ap.add_argument("--pmc-summary", metavar="DIR", help="print the per-kernel table of rocprofv3 counter CSVs")
    ap.add_argument("--pmc-by-grid", action="store_true", help="--pmc-summary: one row per kernel and grid size")
    ap.add_argument("--pmc-segments", metavar="DIR",
                    help="tracer kernels of a --profile-pass run from DIR/pmc_{valu,fetch,write,l2}, room and office")
    ap.add_argument("--kt-segments", metavar="DIR",
                    help="tracer kernels of a --kernel-trace run from DIR, split into the timed steps (--warmup, "
                         "--steps) and the isolated step by the discovery launches, --kt-traces per step")
    ap.add_argument("--kt-traces", type=int, default=3,
                    help="--kt-segments: traces per step (3 for room_boxes_ms, 10 for the office)")
    ap.add_argument("--only-office", action="store_true",
                    help="run only the office workload at its stated size and print its object")
    ap.add_argument("--valu-probe", action="store_true",
                    help="run the VALU issue-rate probe kernels (for a counter run) and exit")
    ap.add_argument("--cpu-sample-fib", type=int, default=1 << 18)
    ap.add_argument("--cpu-sample-ms-fib", type=int, default=1 << 16)
    args = ap.parse_args()
    if args.pmc_segments:
        print(pmc_segments(args.pmc_segments))
        return
    if args.kt_segments:
        print(kt_segments(args.kt_segments, args.warmup, args.steps, args.kt_traces))
        return
    if args.pmc_summary:
        print(pmc_summary(args.pmc_summary, args.pmc_by_grid))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the path shards by Rx tiles "
              "across GPUs by design only)", file=sys.stderr)
        sys.exit(2)

    import torch

    from automs_amd.tracer import TraceScene, trace, trace_many

    dev = torch.device("cuda:0")
    if args.valu_probe:
        import ctypes as C

        from automs_amd import _lib
        blocks = N_CU * 32
        buf = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        for kind in (0, 1, 2, 3):
            for _ in range(3):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                _lib.check(_lib.lib().ams_valu_probe(C.c_void_p(buf.data_ptr()), blocks, 20000, kind, st),
                           "ams_valu_probe")
                e1.record()
                torch.cuda.synchronize()
                print(f"valu probe kind {kind}: {e0.elapsed_time(e1):.3f} ms", flush=True)
        return
    if args.only_surrogate:
        print(json.dumps(surrogate_training()), flush=True)
        return
    if args.only_contract:
        print(json.dumps(contraction_office(iters=20)), flush=True)
        return
    if args.only_office:
        r = office_full(1 if args.profile_pass else args.office_steps, 0 if args.profile_pass else args.office_warmup,
                        cpu=False, profile=args.profile_pass)
        print(json.dumps(r), flush=True)
        return
    if args.profile_pass:
        args.steps, args.warmup, args.no_opt, args.no_cpu_baseline = 1, 0, True, True
        x = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)  # calibration: 1 GiB read + 1 GiB written per launch
        torch.cuda.synchronize()
        del x, y
    traces = workload_traces(args.workload)
    handles = {k: TraceScene(w.scene) for k, w in traces.items()}
    outs = {k: torch.zeros(w.tx.n, w.rx.n, 2, dtype=torch.float32, device=dev) for k, w in traces.items()}

    def step_sequential():
        agg = {}
        for k, w in traces.items():
            outs[k].zero_()
            _, st, _ = trace(handles[k], w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                             w.gridsize, result=outs[k], flags=args.trace_flags)
            for f, v in st.items():
                agg[f] = agg.get(f, 0) + v
        return agg

    def step():
        agg = {}
        if args.sequential:
            sts = []
            for k, w in traces.items():
                outs[k].zero_()
                _, st, _ = trace(handles[k], w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                                 w.gridsize, result=outs[k], flags=args.trace_flags)
                sts.append(st)
        else:
            for k in traces:
                outs[k].zero_()
            res = trace_many([dict(ts=handles[k], tx=w.tx, rx=w.rx, wave_len=w.wave_len,
                                   max_reflection=w.max_reflection, fib_n=w.fib_n, per_element=w.per_element,
                                   gridsize=w.gridsize, result=outs[k], flags=args.trace_flags)
                              for k, w in traces.items()])
            sts = [r[1] for r in res]
        for st in sts:
            for f, v in st.items():
                agg[f] = agg.get(f, 0) + v
        return agg

    for i in range(args.warmup):
        step()
        print(f"warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tot = {}
    for i in range(args.steps):
        st = step()
        for f, v in st.items():
            tot[f] = tot.get(f, 0) + v
        print(f"step {i + 1}/{args.steps}: {st['t_total_ms']:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    queries = tot["fwd_queries"] + tot["bwd_queries"]
    kb = kernel_bytes(tot)
    kt = {"k_discover": tot["t_discover_ms"], "k_evaluate": tot["t_evaluate_ms"]}
    kl = {"k_discover": tot["n_discover_launches"], "k_evaluate": tot["n_evaluate_launches"]}
    # One more step with the traces one after the other (untimed): each
    # kernel's launch durations alone on the GPU.  The largest share of that
    # GPU time names the kernel of the roofline object; concurrent launches
    # share the GPU, so their live durations are longer.
    iso = step_sequential()
    ikt = {"k_discover": iso["t_discover_ms"], "k_evaluate": iso["t_evaluate_ms"]}
    ikl = {"k_discover": iso["n_discover_launches"], "k_evaluate": iso["n_evaluate_launches"]}
    ikb = kernel_bytes(iso)
    dom = max(ikt, key=ikt.get)
    q_of = {"k_discover": "fwd_queries", "k_evaluate": "bwd_queries"}
    lds = args.workload != "office"  # the room scenes are staged in LDS
    roof = valu_roofline(args.workload, dom, tot[q_of[dom]], kl[dom], kt[dom], kb[dom], lds)
    r_iso = valu_roofline(args.workload, dom, iso[q_of[dom]], ikl[dom], ikt[dom], ikb[dom], lds)
    roof["isolated"] = {k: r_iso.get(k) for k in ("avg_launch_ms", "achieved", "frac")}
    # The two kernels' isolated shares are close on the room (about 13.4 ms
    # each per step), so the one named above can change from run to run: the
    # other one's figures are given beside it.
    oth = "k_evaluate" if dom == "k_discover" else "k_discover"
    r_oth = valu_roofline(args.workload, oth, tot[q_of[oth]], kl[oth], kt[oth], kb[oth], lds)
    r_oth_iso = valu_roofline(args.workload, oth, iso[q_of[oth]], ikl[oth], ikt[oth], ikb[oth], lds)
    roof["other_kernel"] = {k: r_oth.get(k) for k in ("kernel", "avg_launch_ms", "achieved", "frac",
                                                       "valu_inst_per_query", "valu_busy_pmc")}
    roof["other_kernel"]["isolated"] = {k: r_oth_iso.get(k) for k in ("avg_launch_ms", "achieved", "frac")}
    roof["isolated_ms_per_step"] = dict(ikt)  # one step
    roof["step"] = step_valu(args.workload, tot, dt)
    roof["streams"] = 1 if args.sequential else len(traces)
    roof["note"] = ("live durations are those of the timed steps, whose traces run concurrently on separate "
                    "streams and share the GPU; isolated = one untimed step with the traces in turn")
    n_channels = sum(w.tx.n * w.rx.n for w in traces.values())
    line = {
        "metric": "Mrays/s & Rx-channels/s; phase-map opt sec to <=0.1 dB of plain optimum",
        "value": queries / dt / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "traces": {k: {"n_tx": w.tx.n, "n_rx": w.rx.n, "tris": w.scene.n_tri, "max_reflection": w.max_reflection,
                           "rays": (w.fib_n * w.tx.n if w.per_element else w.fib_n)} for k, w in traces.items()},
            "parallelism": "single GPU",
        },
        "rx_channels_per_s": n_channels * args.steps / dt,
        "primary_rays_per_s": tot["rays"] / dt,
        "per_step": {k: tot[k] / args.steps for k in ("rays", "paths", "tasks", "rounds", "chains", "successes",
                                                      "fwd_queries", "bwd_queries", "fwd_nodes", "fwd_tris",
                                                      "fwd_spheres", "bwd_nodes", "bwd_tris", "t_discover_ms",
                                                      "t_evaluate_ms", "t_total_ms")},
        "roofline": roof,
    }
    if args.workload == "room_boxes_ms" and not args.no_opt:
        from automs_amd.phasemap import Metasurface, channels_from_traces, time_to_target

        def cplx(k):
            return torch.complex(outs[k][..., 0], outs[k][..., 1])
        ch = channels_from_traces(cplx("H_d"), cplx("G"), cplx("F"))
        ms = Metasurface(None, traces["G"].ms[0]["size"], None, None, [], max_phase=360.0)
        r = time_to_target(ch, ms, steps=args.opt_steps, graph=not args.no_graph)
        print(f"phase map: {r}", file=sys.stderr, flush=True)
        line["phase_map"] = {
            "metric": "seconds to within 0.1 dB of the plain FP32 optimum (mean best-code Rx power)",
            "seconds_to_0p1dB": r["seconds_to_target"], "steps_to_0p1dB": r["mfma_steps_to_target"],
            "adam_steps": args.opt_steps, "mfma_seconds_all_steps": r["mfma_seconds_500"],
            "plain_fp32_seconds_all_steps": r["plain_seconds_500"],
            "plain_optimum_cov_db": r["plain_optimum_cov_db"], "mfma_best_cov_db": r["mfma_best_cov_db"],
            "start_cov_db": r["start_cov_db"], "codes": 30, "objective": "mean log2(1+SNR), N0 = -90 dBm",
            "mfma_step_hip_graph": r["mfma_graph"], "mfma_ms_per_step": r["mfma_ms_per_step"],
            "plain_fp32_ms_per_step": r["plain_ms_per_step"],
            "timing": "HIP events on the stream from before the first step (warm-up and capture included)"}
        if not args.no_cpu_baseline:
            r_cpu = cpu_phase_map(ch, ms.max_phase_rad, r["plain_optimum_cov_db"])
            print(f"phase map host baseline: {r_cpu}", file=sys.stderr, flush=True)
            line["phase_map"]["cpu_baseline"] = r_cpu
    if not args.no_contract:
        r = contraction_office(iters=5 if args.profile_pass else 20)
        print(f"contraction: {r}", file=sys.stderr, flush=True)
        line["contraction_office_batched"] = r
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_sample_fib, args.cpu_sample_ms_fib)
    if not args.no_office and args.workload != "office":
        r = office_full(1 if args.profile_pass else args.office_steps, 0 if args.profile_pass else args.office_warmup,
                        cpu=not args.no_cpu_baseline, profile=args.profile_pass)
        print(f"office: {r}", file=sys.stderr, flush=True)
        line["office_full"] = r
    if not args.no_reference_run and not args.profile_pass:
        r = reference_run()
        print(f"reference run: {r}", file=sys.stderr, flush=True)
        line["reference_run"] = r
    if not args.no_placement_search and not args.profile_pass:
        r = surrogate_training()
        print(f"surrogate training: {r}", file=sys.stderr, flush=True)
        line["surrogate_training"] = r
        r = placement_search(args.sa_steps)
        print(f"placement search: {r}", file=sys.stderr, flush=True)
        line["placement_search"] = r
    if args.profile_pass:
        # every query traced in this process (timed and isolated steps), for
        # the per-query instruction counts of a counter run
        room = {f: tot.get(f, 0) + iso.get(f, 0) for f in ("fwd_queries", "bwd_queries", "fwd_nodes", "bwd_nodes",
                                                         "n_discover_launches", "n_evaluate_launches")}
        line["profile_totals"] = {args.workload: room}
        if "office_full" in line:
            line["profile_totals"]["office"] = line["office_full"]["totals"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
