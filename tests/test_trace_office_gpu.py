"""Tracer parity on the seeded office (100,046 triangles, 5 reflections):
the scene hierarchy is read from global memory (too large for LDS), the Rx
grid goes through the lattice query and the evaluation through its
analytic pre-check.  Reduced sizes (2 x 2 AP, 40 x 28 Rx over the whole
floor, 2^16 rays) keep the FP64 oracle to seconds.  The tie-attributed bars
of tests/_parity.py, as for the room.
"""
import sys

import numpy as np
import pytest

from _parity import check_discovery_fib, check_h, oracle_trace, to_complex

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_office_trace_vs_oracle(gpu):
    import torch

    from automs_amd import scenes
    from automs_amd.tracer import TraceScene, trace
    from oracle import oracle as O
    w = scenes.office(fib_n=1 << 16, n_rx=(40, 28), ap=(2, 2), max_reflection=5)["H_d"]
    _flush(f"case: office H_d {w.scene.n_tri} triangles, K=5, {w.tx.n} Tx x {w.rx.n} Rx, {w.fib_n} rays")
    sc = O.OracleScene.from_scene(w.scene)
    ts = TraceScene(w.scene)
    _, st, gk = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                      keys_cap=1 << 20)
    ok, ost = check_discovery_fib(sc, w, gk)
    # evaluation from the oracle's records, with the neighbour expansion
    Ho, oe, mask = oracle_trace(w, sc, seed=ok)
    res, ge, _ = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                       seed_keys=ok)
    torch.cuda.synchronize()
    assert oe["paths"] > 100
    r = check_h(to_complex(res), Ho, mask, ge, oe, "office H_d:")
    assert r["rel"] < 1e-4


def test_office_device_rounds_match_host_rounds(gpu):
    """Office H_d at its full Rx grid (8 x 8 AP, 256 x 256 Rx, 2^18 rays):
    a (path, tile) map above 2^21 pairs, so the expansion rounds are built on
    the device by default; against the host loop (flags bit 12) the same
    rounds, tasks and successes, and H equal to FP64 accumulation order."""
    import torch

    from automs_amd import scenes
    from automs_amd.tracer import TraceScene, trace
    w = scenes.office(fib_n=1 << 18)["H_d"]
    _flush(f"case: office device-built rounds vs host loop, {w.tx.n} Tx x {w.rx.n} Rx, {w.fib_n} rays")
    ts = TraceScene(w.scene)
    out = {}
    for f in (0, 4096):
        res, st, _ = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize, flags=f)
        torch.cuda.synchronize()
        out[f] = (res.double().cpu().numpy(), st)
    ts.close()
    (Ha, sa), (Hb, sb) = out[0], out[4096]
    tiles = ((256 + 4) // 5) ** 2
    rel = np.linalg.norm(Ha - Hb) / np.linalg.norm(Hb)
    _flush(f"  paths {sa['paths']} x {tiles} tiles; rounds {sa['rounds']} vs {sb['rounds']}, tasks {sa['tasks']} vs "
           f"{sb['tasks']}, successes {sa['successes']} vs {sb['successes']}, H rel {rel:.1e}; evaluation "
           f"{sa['t_evaluate_ms']:.1f} vs {sb['t_evaluate_ms']:.1f} ms, whole trace {sa['t_total_ms']:.1f} vs "
           f"{sb['t_total_ms']:.1f} ms")
    assert sa["paths"] * tiles > (1 << 21)
    assert sa["rounds"] == sb["rounds"] and sa["tasks"] == sb["tasks"]
    assert sa["successes"] == sb["successes"] and sa["chains"] == sb["chains"]
    assert rel < 1e-6
