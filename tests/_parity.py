"""Tie-attributed parity checks of the tracer against the FP64 oracle, shared
by the GPU test modules.

The bars (the HIP path against the oracle on the same inputs):
  * discovery: every (path, tile) record of a ray (Fibonacci mode) or a
    launch thread (reference mode) that meets no near tie is found, and the
    two record sets differ in at most as many records as the oracle counts
    tie rays (tie-thread records in reference mode);
  * evaluation: every (Tx, Rx) entry without a near-tie chain is within 1e-4
    of its row's largest term, the success counts differ by at most the
    counted tie chains, and an Rx whose coverage differs by more than
    0.05 dB holds a tie entry;
near ties being events within TIE_EPS metres of a boundary or of another
event (the oracle's count, FP64).
"""
import sys

import numpy as np

TIE_EPS = 1e-6


def flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def keyset(k):
    return set(map(tuple, np.asarray(k, dtype=np.uint64).tolist()))


def check_discovery_fib(sc, w, gpu_keys):
    """Fibonacci discovery: returns (oracle keys, oracle stats)."""
    from oracle import oracle as O
    args = (sc, w.tx.pos, w.rx.pos, w.rx.radius, w.max_reflection, w.gridsize, w.fib_n, w.per_element)
    with O.tie_bookkeeping(eps=TIE_EPS):
        ok, ost = O.discover(*args)
    with O.tie_bookkeeping(eps=TIE_EPS, skip=True):
        clean, _ = O.discover(*args)
    a, b, c = keyset(ok), keyset(gpu_keys), keyset(clean)
    flush(f"  discovery: oracle {len(a)} gpu {len(b)} symdiff {len(a ^ b)} tie rays {ost['ties']} "
          f"tie-free records missing on the gpu {len(c - b)}")
    assert not (c - b)
    assert len(a ^ b) <= ost["ties"]
    return ok, ost


def check_discovery_ref(sc, w, launch, gpu_keys):
    """Reference-mode discovery (per-thread gating: a tie can change the
    thread's later records)."""
    from oracle import oracle as O
    args = (sc, w.tx.pos, w.rx.pos, w.rx.radius, w.max_reflection, w.gridsize, *launch)
    with O.tie_bookkeeping(eps=TIE_EPS):
        ok, ost = O.discover_ref(*args)
    with O.tie_bookkeeping(eps=TIE_EPS, skip=True):
        clean, cst = O.discover_ref(*args)
    a, b, c = keyset(ok), keyset(gpu_keys), keyset(clean)
    tie_records = ost["raw"] - cst["raw"]
    flush(f"  reference discovery: oracle {len(a)} gpu {len(b)} symdiff {len(a ^ b)} tie threads "
          f"{cst['tie_threads']} ({tie_records} records); tie-free records missing on the gpu {len(c - b)}")
    assert not (c - b)
    assert len(a ^ b) <= tie_records
    return ok, ost


def oracle_trace(w, sc, flags=0, seed=None):
    """Oracle trace with the tie mask: (H, stats, mask [n_tx, n_rx])."""
    from oracle import oracle as O
    with O.tie_bookkeeping(eps=TIE_EPS, n_src=w.tx.n, n_dst=w.rx.n) as tb:
        Ho, ost, _ = O.trace_workload(w, flags=flags, seed=seed, sc=sc)
    return Ho, ost, tb.mask


def to_complex(res):
    Hg = res.double().cpu().numpy()
    return Hg[..., 0] + 1j * Hg[..., 1]


def check_h(Hg, Ho, mask, st=None, ost=None, what=""):
    """Per-entry, success-count and coverage bars; returns the figures."""
    row = np.abs(Ho).max(axis=1, keepdims=True)
    el = np.abs(Hg - Ho) / np.maximum(row, 1e-300)
    bad = el > 1e-4
    co = 20 * np.log10(np.abs(Ho.sum(0)) + 1e-300)
    cg = 20 * np.log10(np.abs(Hg.sum(0)) + 1e-300)
    m = co > -120
    ddb = np.abs(co - cg)
    badc = (ddb > 0.05) & m
    tie_cols = mask.any(0)
    rel = np.linalg.norm(Hg - Ho) / np.linalg.norm(Ho)
    clean_max = float(el[~mask].max()) if (~mask).any() else 0.0
    flush(f"  {what} H rel fro {rel:.2e}; per entry / row max: max {el.max():.2e}, tie-free max {clean_max:.2e}; "
          f"entries > 1e-4: {bad.sum()} ({(bad & ~mask).sum()} tie-free); tie entries {mask.sum()}; "
          f"coverage max {ddb[m].max() if m.any() else 0:.4f} dB, Rx > 0.05 dB: {badc.sum()} "
          f"({(badc & ~tie_cols).sum()} without a tie entry)")
    if (bad & ~mask).any():
        ii, jj = np.nonzero(bad & ~mask)
        for i, j in list(zip(ii, jj))[:8]:
            flush(f"    tie-free entry over the bar: tx {i} rx {j} gpu {Hg[i, j]:.6e} oracle {Ho[i, j]:.6e} "
                  f"err/rowmax {el[i, j]:.2e}")
    assert not (bad & ~mask).any()
    assert not (badc & ~tie_cols).any()
    if st is not None:
        flush(f"  successes gpu {st['successes']} oracle {ost['successes']} tie chains {ost['bwd_ties']}; "
              f"tasks gpu {st['tasks']} oracle {ost['tasks']}")
        assert abs(st["successes"] - ost["successes"]) <= ost["bwd_ties"]
    return dict(rel=rel, el_max=float(el.max()), clean_max=clean_max, n_bad=int(bad.sum()),
                n_badc=int(badc.sum()))
