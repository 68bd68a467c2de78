"""Config-3 scale contraction on traced office channels: 4 metasurface phase
maps x 64 AP beams (M = 256 rows), K = 2304 + 64, N = 65,536 Rx.

Split-bf16 MFMA and plain FP32 against the FP64 oracle (complex128 on the
CPU) on a 4,096-Rx sample of columns; errors are stated per Rx.  Bar: the
median per-Rx relative error < 1e-4 and the 99th percentile < 1e-3 (nulls of
|y| make the worst Rx larger; they are reported, not hidden), and the
mean-log-SNR loss within 1e-4 relative.
"""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _flush(m):
    print(m, flush=True)
    sys.stdout.flush()


def test_office_batched_contraction_vs_fp64(gpu):
    from automs_amd import scenes
    from automs_amd.phasemap import Metasurface, batched_model
    from automs_amd.tracer import TraceScene, trace
    _flush("case: office batched contraction (traces G0, F0, H_d at 2^18 rays)")
    tr = scenes.office(fib_n=1 << 18, ms_fib_n=1 << 18)
    ch = {}
    for key, name in (("src_to_dst", "H_d"), ("src_to_ms{0}", "G0"), ("ms{0}_to_dst", "F0")):
        w = tr[name]
        res, st, _ = trace(TraceScene(w.scene), w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                           w.gridsize)
        ch[key] = torch.complex(res[..., 0], res[..., 1])
        _flush(f"  traced {name}: paths={st['paths']} chains={st['chains']} {st['t_total_ms']:.0f} ms")
    ms = Metasurface(None, (48, 48), None, None, [], max_phase=360.0)
    m_mfma = batched_model(ch, ms)
    m_f32 = batched_model(ch, ms, plain=True)
    with torch.no_grad():
        y1 = m_mfma.received()
        y2 = m_f32.received()
        A = m_mfma.stacked_operand().cpu().to(torch.complex128)
    torch.cuda.synchronize()
    cols = torch.arange(0, 65536, 16)
    Fcat = torch.cat([ch["ms{0}_to_dst"], ch["src_to_dst"]], 0)[:, cols].cpu().to(torch.complex128)
    y64 = A @ Fcat
    keep = y64.abs() > 0
    for name, y in (("mfma split-bf16", y1), ("plain fp32", y2)):
        yc = y[:, cols].cpu().to(torch.complex128)
        err = (yc - y64).abs()
        rel_rx = (err.norm(dim=0) / y64.norm(dim=0).clamp_min(1e-300))[keep.any(0)]
        rel_el = (err[keep] / y64.abs()[keep])
        q = np.quantile(rel_rx.numpy(), [0.5, 0.99, 1.0])
        _flush(f"  {name}: per-Rx rel err median={q[0]:.2e} p99={q[1]:.2e} max={q[2]:.2e}; "
               f"per-element median={rel_el.median():.2e}")
        assert q[0] < 1e-4 and q[1] < 1e-3
    l1 = m_mfma().item()
    l2 = m_f32().item()
    _flush(f"  loss mfma={l1:.6f} f32={l2:.6f}")
    assert abs(l1 - l2) < 1e-4 * abs(l2)
