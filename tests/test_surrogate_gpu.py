"""GPU checks of the placement surrogate: the hand-written conv forward,
input-gradient and weight-gradient kernels, on bf16 MFMA and on the FP32
vector ALU, against torch-CPU float64 autograd on the same bf16-rounded
weights (the MFMA kernels split every FP32 value exactly into three bf16
parts).  Bars: gradients within 1e-4 relative, the forward and the loss
within 1e-5; and a small traced dataset on which training must beat the
mean predictor."""
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as Fn

pytestmark = pytest.mark.gpu


def _flush(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mfma", [True, False])
def test_conv_net_forward_backward_vs_oracle(gpu, mfma):
    from automs_amd.surrogate import SurrogateNet
    _flush(f"case: surrogate conv net fwd/bwd vs float64 autograd (mfma={mfma})")
    net = SurrogateNet(seed=3, mfma=mfma)
    g = torch.Generator().manual_seed(1)
    X = torch.randn(3, 9, 32, 48, generator=g)
    T = torch.randn(3, 1, 32, 48, generator=g)
    pred = net.forward(X.to(gpu), keep=True)
    loss, gpred = net.mse(pred, T.to(gpu))
    net.backward(gpred)
    torch.cuda.synchronize()
    ws = [w.clone().requires_grad_(True) for w in net.weights_f64()]
    a = X.double()
    for l in range(4):
        a = Fn.conv2d(a, ws[2 * l], ws[2 * l + 1], padding=1)
        if l < 3:
            a = torch.relu(a)
    lo = ((a - T.double()) ** 2).mean()
    lo.backward()
    rel_out = ((pred.cpu().double() - a).norm() / a.norm()).item()
    bar = 1e-5
    _flush(f"  output rel err {rel_out:.2e}, loss rel err {abs(loss.item() - lo.item()) / lo.item():.2e}")
    assert rel_out < bar, rel_out
    assert abs(loss.item() - lo.item()) < bar * lo.item()
    for i, w in enumerate(ws):
        gg = net.grad_of(i).cpu().double()
        rel = ((gg - w.grad).norm() / w.grad.norm()).item()
        _flush(f"  param {i} {tuple(w.shape)} grad rel err {rel:.2e}")
        assert rel < 1e-4, (i, rel)


def test_surrogate_learns_traced_gain_maps(gpu):
    from automs_amd import surrogate as S
    _flush("case: surrogate training on a small traced dataset")
    P, Sid, Y = S.generate_dataset(2, 96, seed=5, fib_ap=1 << 16, fib_ms=1 << 16, log=_flush, label="proxy")
    X = S.build_inputs(P, Sid)
    Y = Y.reshape(-1, 1, 64, 64)
    Xtr, Ytr, Xva, Yva = X[:160], Y[:160], X[160:], Y[160:]
    net = S.SurrogateNet(seed=0)
    hist = S.train(net, Xtr, Ytr, epochs=40, batch=32, lr=3e-3, scale=10.0)
    with torch.no_grad():
        pv = net.forward(Xva) * 10
    mse = float(((pv - Yva) ** 2).mean())
    base = float(((Ytr.mean() - Yva) ** 2).mean())
    _flush(f"  train mse {hist[0] * 100:.2f} -> {hist[-1] * 100:.2f} dB^2; val mse {mse:.2f} vs mean predictor {base:.2f}")
    assert hist[-1] < 0.5 * hist[0]
    assert mse < 0.8 * base
    # what the search gains from it: held-out placements ranked by predicted
    # score against their traced score (the search without the network can
    # only pick at random, i.e. the mean)
    sp = pv.mean((1, 2, 3)).cpu().numpy()
    st = Yva.mean((1, 2, 3)).cpu().numpy()
    rank = lambda x: np.argsort(np.argsort(x))  # noqa: E731
    rho = float(np.corrcoef(rank(sp), rank(st))[0, 1])
    top = np.argsort(-sp)[:8]
    _flush(f"  placement score: Spearman {rho:.2f}; traced score of the 8 best predicted {st[top].mean():.2f} dB "
           f"vs all {st.mean():.2f} dB")
    assert rho > 0.5
    assert st[top].mean() > st.mean()
