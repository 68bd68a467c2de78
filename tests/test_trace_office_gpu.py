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
