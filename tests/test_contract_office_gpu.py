"""Config-3 scale contraction on traced office channels: 4 metasurface phase
maps x 64 AP beams (M = 256 rows), K = 2304 + 64, N = 65,536 Rx.

Split-bf16 MFMA and plain FP32 against the FP64 oracle (complex128 on the
CPU) on a 4,096-Rx sample of columns; errors are stated per Rx.  Bar: every
sampled Rx within 1e-4 relative (its error norm over the 256 rows against
its own norm; Rx with y = 0, nulls, are counted and reported), and the
mean-log-SNR loss within 1e-4 relative.

The hand-written backward at the same shape (M = 256 rows, K = 2304
metasurface elements + 64 AP elements) on a sample of 8,192 Rx columns:
the phase-map gradient within 1e-4 of the complex128 autograd gradient, and
within 1e-3 of central finite differences (h = 1e-3 rad) of the complex128
loss on the 16 entries with the largest gradient.
"""
import math
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _flush(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def office_channels(gpu):
    from automs_amd import scenes
    from automs_amd.tracer import TraceScene, trace
    _flush("case: office channels (traces G0, F0, H_d at 2^18 rays)")
    tr = scenes.office(fib_n=1 << 18, ms_fib_n=1 << 18)
    ch = {}
    for key, name in (("src_to_dst", "H_d"), ("src_to_ms{0}", "G0"), ("ms{0}_to_dst", "F0")):
        w = tr[name]
        res, st, _ = trace(TraceScene(w.scene), w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                           w.gridsize)
        ch[key] = torch.complex(res[..., 0], res[..., 1])
        _flush(f"  traced {name}: paths={st['paths']} chains={st['chains']} {st['t_total_ms']:.0f} ms")
    return ch


def test_office_batched_gradient_vs_fp64(gpu, office_channels):
    """Config-3 shape backward (64 beams x 4 maps, K = 2304 + 64) on 8192 Rx
    columns: the Function's gradient of the 4 x 2304 phases against
    complex128 autograd and central finite differences."""
    from automs_amd.phasemap import Metasurface, batched_model
    from oracle import contract_ref as R
    ch = office_channels
    cols = torch.arange(0, 65536, 8)
    sub = {"src_to_dst": ch["src_to_dst"][:, cols].contiguous(), "src_to_ms{0}": ch["src_to_ms{0}"],
           "ms{0}_to_dst": ch["ms{0}_to_dst"][:, cols].contiguous()}
    ms = Metasurface(None, (48, 48), None, None, [], max_phase=360.0)
    model = batched_model(sub, ms)
    loss = model()
    loss.backward()
    g_gpu = model.ms_theta[0].grad.detach().cpu().double()
    th = model.ms_theta[0].detach().cpu().double().requires_grad_(True)
    Hd = sub["src_to_dst"].cpu().to(torch.complex128)
    G = sub["src_to_ms{0}"].cpu().to(torch.complex128)
    F = sub["ms{0}_to_dst"].cpu().to(torch.complex128)
    A, ta = model.antenna_A.cpu(), model.antenna_theta.detach().cpu().double()

    def loss64(t):
        return R.loss_logsnr(R.received(A, ta, Hd, [G], [F], [t], [2 * math.pi]), model.n_maps, model.n0)
    lo = loss64(th)
    lo.backward()
    g64 = th.grad
    rel = float((g_gpu - g64).norm() / g64.norm())
    _flush(f"  loss gpu {loss.item():.6f} complex128 {lo.item():.6f}; gradient rel err {rel:.2e} "
           f"(|g| = {g64.norm():.3e}, {g64.numel()} entries)")
    assert abs(loss.item() - lo.item()) < 1e-4 * abs(lo.item())
    assert rel < 1e-4
    # central differences of the complex128 loss on the largest entries; the
    # perturbed received signal is a rank-one update of the exact one (y is
    # linear in each element's u)
    with torch.no_grad():
        W = A.double() * torch.exp(-1j * ta)
        WG = W @ G
        Y = R.received(A, ta, Hd, [G], [F], [th.detach()], [2 * math.pi])
        h = 1e-3
        idx = torch.topk(g64.abs().reshape(-1), 16).indices
        worst = 0.0
        for flat in idx.tolist():
            m, k = divmod(flat, th.shape[1])
            vals = []
            for s_ in (1, -1):
                du = torch.exp(1j * (th[m, k].detach() + s_ * h)) - torch.exp(1j * th[m, k].detach())
                Yp = Y.clone()
                rows = slice(m * model.code_num, (m + 1) * model.code_num)
                Yp[rows] += (WG[:, k] * du)[:, None] * F[k][None, :]
                vals.append(R.loss_logsnr(Yp, model.n_maps, model.n0).item())
            fd = (vals[0] - vals[1]) / (2 * h)
            err = abs(fd - g_gpu[m, k].item()) / abs(fd)
            worst = max(worst, err)
        _flush(f"  central differences (h = {h}) on the 16 largest entries: worst rel err {worst:.2e}")
        assert worst < 1e-3


def test_office_batched_contraction_vs_fp64(gpu, office_channels):
    from automs_amd.phasemap import Metasurface, batched_model
    ch = office_channels
    _flush("case: office batched contraction M = 256, K = 2368, N = 65,536")
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
        _flush(f"  {name}: per-Rx rel err median={q[0]:.2e} p99={q[1]:.2e} max={q[2]:.2e} "
               f"({int((~keep.any(0)).sum())} null Rx of {len(cols)}); per-element median={rel_el.median():.2e}")
        assert q[2] < 1e-4
    l1 = m_mfma().item()
    l2 = m_f32().item()
    _flush(f"  loss mfma={l1:.6f} f32={l2:.6f}")
    assert abs(l1 - l2) < 1e-4 * abs(l2)
