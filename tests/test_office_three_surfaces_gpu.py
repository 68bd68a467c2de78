"""Config 2's three metasurfaces on traced office channels (8 x 8 AP, three
48 x 48 surfaces, the ms-to-ms channels T01, T02, T12, 2^18 rays per
trace): the chain kernels' loss and phase gradients with chains of up to
two surfaces (go.py's default max_ms_reflex) against complex128 autograd
through go.py's literal chain enumeration, on a 2,048-Rx sample of columns,
each within 1e-4 relative; then a short optimisation over all 65,536 Rx
replayed from a HIP graph must raise the coverage."""
import math

import pytest
import torch

from _parity import flush

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def office_all(gpu):
    from automs_amd import scenes
    from automs_amd.tracer import TraceScene, trace_many
    flush("case: office channels, all ten traces at 2^18 rays")
    tr = scenes.office(fib_n=1 << 18, ms_fib_n=1 << 18)
    handles = {k: TraceScene(w.scene) for k, w in tr.items()}
    jobs = [dict(ts=handles[k], tx=w.tx, rx=w.rx, wave_len=w.wave_len, max_reflection=w.max_reflection,
                 fib_n=w.fib_n, per_element=w.per_element, gridsize=w.gridsize) for k, w in tr.items()]
    res = trace_many(jobs)
    ch = {}
    for (k, w), (r, st, _) in zip(tr.items(), res):
        ch[scenes.OFFICE_CHANNEL_KEYS[k]] = torch.complex(r[..., 0], r[..., 1])
        flush(f"  {k}: {w.tx.n} x {w.rx.n}, paths {st['paths']}, successes {st['successes']}")
    for h in handles.values():
        h.close()
    return ch


def test_office_three_surfaces_vs_fp64(gpu, office_all):
    from automs_amd.phasemap import Metasurface, PhaseMapModel
    from oracle import contract_ref as R
    cols = torch.arange(0, 65536, 32)
    sub = {k: (v[:, cols].contiguous() if k.endswith("_dst") else v) for k, v in office_all.items()}
    for k in ("ms{0}_to_ms{1}", "ms{0}_to_ms{2}", "ms{1}_to_ms{2}"):
        assert sub[k].abs().max() > 0, k
    ms = [Metasurface(None, (48, 48), None, None, [], max_phase=300.0) for _ in range(3)]
    g = torch.Generator().manual_seed(5)
    th = [torch.rand(1, 2304, generator=g) * 2 * math.pi for _ in range(3)]
    model = PhaseMapModel("office3", sub, ms, code_num=8, objective="logsnr", ms_theta=th, max_ms_reflex=2)
    assert model.chain_operand and model.chain is not None
    loss = model()
    loss.backward()
    ch64 = {k: v.cpu().to(torch.complex128) for k, v in sub.items()}
    th_ap = model.antenna_theta.detach().cpu().double().requires_grad_(True)
    th64 = [t.double().reshape(-1).requires_grad_(True) for t in th]
    Y = R.received_chains(model.antenna_A.cpu(), th_ap, ch64, th64, [300 / 180 * math.pi] * 3, 2)
    lo = R.loss_logsnr(Y, 1, model.n0)
    lo.backward()
    Yg = model.received().detach().cpu().to(torch.complex128)
    per_rx = ((Yg - Y.detach()).abs().norm(dim=0) / Y.detach().abs().norm(dim=0).clamp_min(1e-300))
    rel_l = abs(loss.item() - lo.item()) / abs(lo.item())
    rel_ap = ((model.antenna_theta.grad.cpu().double() - th_ap.grad).norm() / th_ap.grad.norm()).item()
    rel_ms = [((model.ms_theta[i].grad.cpu().double().reshape(-1) - th64[i].grad).norm()
               / th64[i].grad.norm()).item() for i in range(3)]
    flush(f"  per-Rx error: median {per_rx.median():.2e} max {per_rx.max():.2e} over {per_rx.numel()} Rx; "
          f"loss rel {rel_l:.1e}; gradient rel AP {rel_ap:.1e}, MS {', '.join(f'{r:.1e}' for r in rel_ms)}")
    assert per_rx.max() < 1e-4
    assert rel_l < 1e-4 and rel_ap < 1e-4 and max(rel_ms) < 1e-4


def test_office_three_surfaces_optimisation(gpu, office_all):
    from automs_amd.phasemap import Metasurface, PhaseMapModel, optimize_model
    ms = [Metasurface(None, (48, 48), None, None, [], max_phase=360.0) for _ in range(3)]
    model = PhaseMapModel("office3", office_all, ms, code_num=30, objective="logsnr", max_ms_reflex=2)
    c0 = float(model.coverage_db()[0])
    optimize_model(model, lr=0.3, step=30, graph=True)
    c1 = float(model.coverage_db()[0])
    flush(f"  coverage {c0:.3f} -> {c1:.3f} dB after 30 steps; losses {model.loss_list[0]:.4f} -> "
          f"{min(model.loss_list):.4f}")
    assert min(model.loss_list) < model.loss_list[0]
    assert c1 > c0
    del model
    torch.cuda.empty_cache()
