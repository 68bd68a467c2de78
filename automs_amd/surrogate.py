"""Learned placement surrogate and the annealing search it steers.

The reference tunes the metasurface placement (wall, offset) with NNI's
simulated-annealing tuner, every trial retracing and re-optimising from
scratch (physical_exp_nni_main.py:12-29, physical_exp_nni_model.py:59-70).
Here an in-process annealing loop searches placements, a small conv net
predicts each candidate's per-Rx gain map, and the ray tracer re-scores every
accepted candidate (the tracer's score is the one that decides).

Network: 4 conv3x3 layers, channels 9 -> 32 -> 32 -> 32 -> 1, ReLU between,
bf16 weights with FP32 master copies, all forward/backward math in
libams.so (ams_conv3x3_fwd / ams_conv3x3_bwd_weight / ams_relu_bwd /
ams_mse), Adam via ams_adam_step.

Score of a placement (PlacementTracer.trial_score): what a trial of
physical_exp_nni_model.py:96-149 reports for it, the coverage after optimising
the AP codebook and the metasurface phases together (30 codes, Adam lr 0.3,
the mean-log-SNR objective, best parameters kept): the mean over the Rx grid
of the best code's power in dB, minus the same figure for the optimised
codebook without a metasurface.  The network's label is that difference per
Rx (trial_gain_map, a 64 x 64 map whose mean is the score).

A cheaper proxy, the per-Rx gain a metasurface with phases matched to every Rx
at once could add to the AP's broadside beam,
    h_d = sum_tx H_d[tx, rx],  g_k = sum_tx G[tx, k],  m = sum_k |g_k| |F[k, rx]|,
    gain = 20 log10((|h_d| + m + eps) / (|h_d| + eps)), clamped to [0, 40] dB
(gain_map), is an upper bound no single phase map reaches and does not rank
placements like the trial (rank_check measures both against it).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .geometry import gen_ms
from .phasemap import Metasurface, PhaseMapModel, optimize_model
from .scenes import ArraySpec, SplitMix64, Workload, room_boxes, room_boxes_scene, WAVE_LEN_60GHZ
from .tracer import TraceScene, trace

ROOM = (10.0, 8.0, 3.0)
GRID = 64
LABEL_SCALE = 1.0  # the network predicts gain / LABEL_SCALE (dB)
EPS = 1e-7
MS_SIZE = (32, 32)


def _p(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _s() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


# ----------------------------------------------------------------- network --

class SurrogateNet:
    """Conv net on the GPU with hand-written kernels: bf16 MFMA implicit
    GEMMs (mfma=True, the default) or the FP32 vector-ALU kernels."""

    def __init__(self, channels=(9, 32, 32, 32, 1), seed: int = 0, device="cuda", mfma: bool = True):
        self.mfma = mfma
        self.ch = list(channels)
        self.dev = torch.device(device)
        g = torch.Generator().manual_seed(seed)
        shapes = []
        for i in range(len(self.ch) - 1):
            shapes.append((self.ch[i + 1], self.ch[i], 3, 3))
            shapes.append((self.ch[i + 1],))
        n = sum(int(np.prod(s)) for s in shapes)
        self.flat = torch.zeros(n, dtype=torch.float32)
        off = 0
        self.views = []
        for s in shapes:
            k = int(np.prod(s))
            if len(s) == 4:  # He initialisation
                fan_in = s[1] * 9
                self.flat[off:off + k] = torch.randn(k, generator=g) * math.sqrt(2.0 / fan_in)
            self.views.append((off, s))
            off += k
        self.flat = self.flat.to(self.dev)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.t = 0
        self.refresh()

    def save(self, path: str) -> None:
        np.save(path, self.flat.detach().cpu().numpy())

    def load(self, path: str) -> None:
        w = torch.from_numpy(np.load(path)).to(self.dev)
        if w.shape != self.flat.shape:
            raise ValueError(f"{path}: {tuple(w.shape)} parameters, the network has {tuple(self.flat.shape)}")
        self.flat.copy_(w)
        self.refresh()

    def param(self, i: int) -> torch.Tensor:
        off, s = self.views[i]
        return self.flat[off:off + int(np.prod(s))].view(s)

    def grad_of(self, i: int) -> torch.Tensor:
        off, s = self.views[i]
        return self.grad[off:off + int(np.prod(s))].view(s)

    def _bf16(self, x: torch.Tensor) -> torch.Tensor:
        x = x.contiguous()
        out = torch.empty(x.numel(), dtype=torch.int16, device=self.dev)
        _lib.check(_lib.lib().ams_to_bf16(_p(x), _p(out), x.numel(), _s()), "ams_to_bf16")
        return out

    def refresh(self):
        """bf16 weight copies from the FP32 masters, and the flipped /
        channel-transposed copies the input-gradient pass uses."""
        L = _lib.lib()
        self.w_bf, self.wt_bf, self.w_frag, self.wt_frag = [], [], [], []
        for l in range(len(self.ch) - 1):
            W = self.param(2 * l)
            self.w_bf.append(self._bf16(W))
            self.wt_bf.append(self._bf16(W.flip(2, 3).transpose(0, 1).contiguous()))
            if self.mfma:
                for src, co, ci, dst in ((self.w_bf[-1], self.ch[l + 1], self.ch[l], self.w_frag),
                                         (self.wt_bf[-1], self.ch[l], self.ch[l + 1], self.wt_frag)):
                    f = torch.empty(L.ams_conv3x3_wfrag_bytes(co, ci), dtype=torch.uint8, device=self.dev)
                    _lib.check(L.ams_conv3x3_pack_weights(_p(src), co, ci, _p(f), _s()), "ams_conv3x3_pack_weights")
                    dst.append(f)

    def weights_f64(self) -> List[torch.Tensor]:
        """The bf16-rounded weights the kernels use, as float64 CPU tensors."""
        out = []
        for l in range(len(self.ch) - 1):
            W = self.param(2 * l).detach().float().cpu().contiguous()
            u = W.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
            u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).to(torch.int64)
            u = torch.where(u >= 2 ** 31, u - 2 ** 32, u).to(torch.int32)
            out.append(u.view(torch.float32).double())
            out.append(self.param(2 * l + 1).detach().cpu().double())
        return out

    def forward(self, X: torch.Tensor, keep: bool = False) -> torch.Tensor:
        L = _lib.lib()
        B, _, H, W = X.shape
        acts = [X.contiguous()]
        a = acts[0]
        nl = len(self.ch) - 1
        for l in range(nl):
            y = torch.empty(B, self.ch[l + 1], H, W, dtype=torch.float32, device=self.dev)
            self._conv(a, l, False, _p(self.param(2 * l + 1)), y, B, H, W, 1 if l < nl - 1 else 0)
            acts.append(y)
            a = y
        if keep:
            self.acts = acts
        return a

    def _conv(self, x: torch.Tensor, l: int, transposed: bool, bias, y: torch.Tensor, B: int, H: int, W: int,
              relu: int):
        """Layer l's convolution (transposed: its input-gradient pass)."""
        L = _lib.lib()
        ci, co = (self.ch[l + 1], self.ch[l]) if transposed else (self.ch[l], self.ch[l + 1])
        if self.mfma:
            wf = self.wt_frag[l] if transposed else self.w_frag[l]
            _lib.check(L.ams_conv3x3_fwd_mfma(_p(x), _p(wf), bias, _p(y), B, ci, co, H, W, relu, _s()),
                       "ams_conv3x3_fwd_mfma")
        else:
            wb = self.wt_bf[l] if transposed else self.w_bf[l]
            _lib.check(L.ams_conv3x3_fwd(_p(x), _p(wb), bias, _p(y), B, ci, co, H, W, relu, _s()), "ams_conv3x3_fwd")

    def backward(self, gY: torch.Tensor):
        """Fills self.grad from dL/dY of the last forward(keep=True)."""
        L = _lib.lib()
        acts = self.acts
        B, _, H, W = gY.shape
        g = gY.contiguous()
        nl = len(self.ch) - 1
        for l in reversed(range(nl)):
            if self.mfma:
                ws = torch.empty(L.ams_conv3x3_dw_workspace_bytes(B, self.ch[l], self.ch[l + 1], H, W),
                                 dtype=torch.uint8, device=self.dev)
                _lib.check(L.ams_conv3x3_bwd_weight_mfma(_p(acts[l]), _p(g), _p(self.grad_of(2 * l)),
                                                         _p(self.grad_of(2 * l + 1)), B, self.ch[l], self.ch[l + 1],
                                                         H, W, _p(ws), _s()), "ams_conv3x3_bwd_weight_mfma")
            else:
                _lib.check(L.ams_conv3x3_bwd_weight(_p(acts[l]), _p(g), _p(self.grad_of(2 * l)),
                                                    _p(self.grad_of(2 * l + 1)), B, self.ch[l], self.ch[l + 1], H, W,
                                                    _s()), "ams_conv3x3_bwd_weight")
            if l == 0:
                break
            gx = torch.empty(B, self.ch[l], H, W, dtype=torch.float32, device=self.dev)
            self._conv(g, l, True, None, gx, B, H, W, 0)
            _lib.check(L.ams_relu_bwd(_p(acts[l]), _p(gx), gx.numel(), _s()), "ams_relu_bwd")
            g = gx

    def mse(self, pred: torch.Tensor, target: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        L = _lib.lib()
        n = pred.numel()
        grad = torch.empty_like(pred)
        part = torch.empty(L.ams_mse_partials(n), dtype=torch.float32, device=self.dev)
        _lib.check(L.ams_mse(_p(pred), _p(target.contiguous()), None, n, _p(grad), _p(part), _s()), "ams_mse")
        return part.sum() / n, grad

    def adam_step(self, lr: float = 2e-3):
        self.t += 1
        _lib.check(_lib.lib().ams_adam_step(_p(self.flat), _p(self.grad), _p(self.m), _p(self.v), self.flat.numel(),
                                            lr, 0.9, 0.999, 1e-8, self.t, _s()), "ams_adam_step")
        self.refresh()


# --------------------------------------------------------------- placements --

@dataclass
class Placement:
    wall: int   # 0: x = 0 (normal +x), 1: x = 10 (-x), 2: y = 0 (+y), 3: y = 8 (-y)
    s: float    # position along the wall, 0..1
    z: float    # centre height (m)

    def pose(self):
        X, Y, _ = ROOM
        if self.wall == 0:
            return 0.0, self.s * Y, self.z, 0.0
        if self.wall == 1:
            return X, self.s * Y, self.z, math.pi
        if self.wall == 2:
            return self.s * X, 0.0, self.z, math.pi / 2
        return self.s * X, Y, self.z, -math.pi / 2

    def ms(self, size=MS_SIZE):
        x, y, z, yaw = self.pose()
        k = [math.cos(yaw), math.sin(yaw), 0.0]
        pos, bh, nrm = gen_ms([x, y, z], k, [0.0, 0.0, -1.0], size, WAVE_LEN_60GHZ / 2)
        return pos.numpy().astype(np.float32), bh, nrm.numpy().astype(np.float32)


def random_placement(rng: SplitMix64) -> Placement:
    return Placement(int(rng.uniform(0, 4)) % 4, rng.uniform(0.1, 0.9), rng.uniform(0.6, 2.4))


def neighbour(p: Placement, rng: SplitMix64) -> Placement:
    if rng.uniform() < 0.1:
        return Placement(int(rng.uniform(0, 4)) % 4, rng.uniform(0.1, 0.9), rng.uniform(0.6, 2.4))
    u1, u2 = max(rng.uniform(), 1e-12), rng.uniform()
    n1 = math.sqrt(-2 * math.log(u1)) * math.cos(2 * math.pi * u2)
    n2 = math.sqrt(-2 * math.log(u1)) * math.sin(2 * math.pi * u2)
    return Placement(p.wall, min(0.9, max(0.1, p.s + 0.06 * n1)), min(2.4, max(0.6, p.z + 0.2 * n2)))


def occupancy(scene_seed: int) -> np.ndarray:
    """Top view of the room at the Rx grid: height of the tallest box over
    each Rx position / room height."""
    sc = room_boxes_scene(scene_seed)
    boxes = sc.tri[12:].reshape(-1, 12, 3, 3)
    xs = np.linspace(0.1, 9.9, GRID)
    ys = np.linspace(0.1, 7.9, GRID)
    occ = np.zeros((GRID, GRID), np.float32)
    for b in boxes:
        lo, hi = b.reshape(-1, 3).min(0), b.reshape(-1, 3).max(0)
        ix = (xs >= lo[0]) & (xs <= hi[0])
        iy = (ys >= lo[1]) & (ys <= hi[1])
        occ[np.ix_(ix, iy)] = np.maximum(occ[np.ix_(ix, iy)], hi[2] / ROOM[2])
    return occ


def features(occ: np.ndarray, p: Placement) -> np.ndarray:
    """9 input channels [9, 64, 64]: occupancy, Rx x / y, offsets to the
    metasurface, distance, cosine to its normal, its height and size."""
    x, y, z, yaw = p.pose()
    xs = np.linspace(0.1, 9.9, GRID)[:, None] * np.ones((1, GRID))
    ys = np.ones((GRID, 1)) * np.linspace(0.1, 7.9, GRID)[None, :]
    dx, dy, dz = xs - x, ys - y, 1.2 - z
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    cosn = (dx * math.cos(yaw) + dy * math.sin(yaw)) / np.maximum(r, 1e-6)
    f = np.stack([occ, xs / 10, ys / 8, dx / 10, dy / 8, r / 10, cosn, np.full_like(xs, z / 3),
                  np.full_like(xs, MS_SIZE[0] * WAVE_LEN_60GHZ / 2 / 0.1)]).astype(np.float32)
    return f


# ------------------------------------------------------------------ tracing --

class PlacementTracer:
    """Traces placements in one scene: H_d once, then G and F per placement
    (reduced ray counts; the scene handle is reused).  gain_map gives the
    network's training target, trial_score the optimised coverage gain that
    a trial reports."""

    def __init__(self, scene_seed: int, fib_ap: int = 1 << 18, fib_ms: int = 1 << 18, max_reflection: int = 3,
                 trial_steps: int = 40, code_num: int = 30):
        self.base = room_boxes(scene_seed, fib_n=fib_ap, max_reflection=max_reflection)
        self.ts = TraceScene(self.base.scene)
        self.fib_ap, self.fib_ms = fib_ap, fib_ms
        w = self.base
        res, _, _ = trace(self.ts, w.tx, w.rx, w.wave_len, w.max_reflection, fib_ap, True, w.gridsize)
        self.Hd = torch.complex(res[..., 0], res[..., 1])
        self.hd = self.Hd.sum(0)  # broadside beam
        self.trial_steps, self.code_num = trial_steps, code_num
        self._base_cov = None
        self._trials = 0
        self.n_traces = 0
        self.gpu_seconds = 0.0

    def channels(self, p: Placement):
        """G [n_ap, K] and F [K, n_rx] of a placement, complex."""
        w = self.base
        pos, _, nrm = p.ms()
        resG, _, _ = trace(self.ts, w.tx, ArraySpec(pos, nrm, 1, 0.06), w.wave_len, w.max_reflection, self.fib_ap,
                           True, w.gridsize)
        resF, _, _ = trace(self.ts, ArraySpec(pos, nrm, 1, 0.5e-5), w.rx, w.wave_len, w.max_reflection, self.fib_ms,
                           False, w.gridsize)
        return torch.complex(resG[..., 0], resG[..., 1]), torch.complex(resF[..., 0], resF[..., 1])

    def gain_map(self, p: Placement) -> torch.Tensor:
        t0 = time.perf_counter()
        G, F = self.channels(p)
        g = G.sum(0)
        m = (g.abs()[:, None] * F.abs()).sum(0)
        a = self.hd.abs()
        gain = (20 * torch.log10((a + m + EPS) / (a + EPS))).clamp(0, 40).reshape(GRID, GRID)
        torch.cuda.synchronize()
        self.gpu_seconds += time.perf_counter() - t0
        self.n_traces += 1
        return gain

    @torch.no_grad()
    def _power_db(self, model: PhaseMapModel) -> torch.Tensor:
        Y = model.received()
        return 10 * torch.log10((Y.real ** 2 + Y.imag ** 2).amax(0) + 1e-30)

    def _optimised_power_db(self, ch: dict, ms_list) -> torch.Tensor:
        """Per-Rx best-code power (dB) after optimising the codebook (and the
        metasurface phases) for trial_steps Adam steps."""
        model = PhaseMapModel("trial", ch, ms_list, code_num=self.code_num, objective="logsnr")
        optimize_model(model, lr=0.3, step=self.trial_steps, graph=True, record=False)
        pw = self._power_db(model)
        del model
        # every trial's graph leaves its private memory pool in torch's cache
        # once freed; hand those pools back to the device now and then, so the
        # tracer's own allocations (outside torch) keep finding memory
        self._trials += 1
        if self._trials % 16 == 0:
            torch.cuda.empty_cache()
        return pw

    def base_power_db(self) -> torch.Tensor:
        """Per-Rx power of the optimised codebook without a metasurface."""
        if self._base_cov is None:
            self._base_cov = self._optimised_power_db({"src_to_dst": self.Hd}, [])
        return self._base_cov

    def trial_gain_map(self, p: Placement) -> torch.Tensor:
        """Per-Rx coverage gain (dB, [64, 64]) of the placement after its
        trial: the AP codebook and the metasurface phases optimised together
        (physical_exp_nni_model.py:96-149 with the log-SNR objective) against
        the optimised codebook alone.  H_d is traced once per scene, without
        the metasurface's footprint (a 16 cm square of wall)."""
        t0 = time.perf_counter()
        base = self.base_power_db()
        G, F = self.channels(p)
        ms = Metasurface(None, MS_SIZE, None, None, [], max_phase=360.0)
        pw = self._optimised_power_db({"src_to_dst": self.Hd, "src_to_ms{0}": G, "ms{0}_to_dst": F}, [ms])
        gain = (pw - base).reshape(GRID, GRID)
        torch.cuda.synchronize()
        self.gpu_seconds += time.perf_counter() - t0
        self.n_traces += 1
        return gain

    def trial_score(self, p: Placement) -> float:
        """The trial's coverage gain (dB): the mean of trial_gain_map."""
        return float(self.trial_gain_map(p).mean())


def generate_dataset(n_scenes: int, per_scene: int, seed: int = 0, fib_ap: int = 1 << 18, fib_ms: int = 1 << 18,
                     log=None, first_scene: int = 0, label: str = "trial"):
    """Gain maps of random placements (label "trial": trial_gain_map;
    "proxy": gain_map): returns (placements [N, 3] (wall, s, z), scene seeds
    [N], gains [N, 64, 64] f32 on the GPU)."""
    rng = SplitMix64(seed * 1000003 + first_scene)
    P, S_, Ys = [], [], []
    for sidx in range(first_scene, first_scene + n_scenes):
        pt = PlacementTracer(sidx, fib_ap, fib_ms)
        for _ in range(per_scene):
            p = random_placement(rng)
            P.append([p.wall, p.s, p.z])
            S_.append(sidx)
            Ys.append(pt.trial_gain_map(p) if label == "trial" else pt.gain_map(p))
        if log:
            log(f"scene {sidx + 1 - first_scene}/{n_scenes}: {pt.n_traces} placements traced in {pt.gpu_seconds:.1f} s")
        del pt
    return np.asarray(P, np.float32), np.asarray(S_, np.int32), torch.stack(Ys).float()


def build_inputs(placements: np.ndarray, scene_ids: np.ndarray) -> torch.Tensor:
    """[N, 9, 64, 64] network inputs on the GPU."""
    occ = {int(s): occupancy(int(s)) for s in np.unique(scene_ids)}
    X = np.stack([features(occ[int(s)], Placement(int(p[0]), float(p[1]), float(p[2])))
                  for p, s in zip(placements, scene_ids)])
    return torch.from_numpy(X).cuda()


def save_dataset(path: str, placements, scene_ids, gains: torch.Tensor) -> None:
    """Gains as float16 (the trial gains are small and signed)."""
    np.savez_compressed(path, placements=placements, scene_ids=scene_ids,
                        gains_f16=gains.cpu().numpy().astype(np.float16))


def load_dataset(paths: List[str]):
    P, S_, G = [], [], []
    for p in paths:
        d = np.load(p)
        P.append(d["placements"])
        S_.append(d["scene_ids"])
        G.append(d["gains_f16"].astype(np.float32) if "gains_f16" in d else d["gains_u8"] * np.float32(40 / 255))
    return np.concatenate(P), np.concatenate(S_), torch.from_numpy(np.concatenate(G)).cuda()


def train(net: SurrogateNet, X: torch.Tensor, Y: torch.Tensor, epochs: int = 20, batch: int = 64, lr: float = 2e-3,
          seed: int = 0, log=None, scale: float = LABEL_SCALE) -> List[float]:
    """Minibatch Adam on the MSE of the gain map (divided by scale)."""
    g = torch.Generator().manual_seed(seed)
    n = X.shape[0]
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g)
        tot = torch.zeros((), device=X.device)
        nb = 0
        lr_ep = lr * (0.5 * (1 + math.cos(math.pi * ep / epochs)))
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch].to(X.device)
            pred = net.forward(X[idx], keep=True)
            loss, gpred = net.mse(pred, Y[idx] / scale)
            net.backward(gpred)
            net.adam_step(lr_ep)
            tot += loss.detach()
            nb += 1
        hist.append(float(tot / max(nb, 1)))
        if log:
            log(f"epoch {ep + 1}/{epochs}: mse {hist[-1]:.4f} (gain/10)^2")
    return hist


def predict_score(net: SurrogateNet, occ: np.ndarray, ps: List[Placement]) -> torch.Tensor:
    X = torch.from_numpy(np.stack([features(occ, p) for p in ps])).cuda()
    return net.forward(X).mean((1, 2, 3)) * LABEL_SCALE


# ---------------------------------------------------------------- annealing --

def anneal(score_fn, steps: int, seed: int, t0: float = 1.0, t1: float = 0.01, rescore=None, start=None,
           stats: Optional[dict] = None):
    """Simulated annealing maximising a placement score.

    Plain search (rescore None): score_fn is the traced score of a candidate.
    Surrogate search: score_fn is the predicted score; the Metropolis test
    compares predictions, and every accepted candidate is re-scored by
    rescore (the tracer); the best placement is chosen on traced scores."""
    rng = SplitMix64(seed)
    cur = start or random_placement(rng)
    cur_val = score_fn(cur)
    best = cur
    best_val = rescore(cur) if rescore else cur_val
    accepted = 0
    if stats is not None:
        stats.update(uphill=[0, 0, 0, 0], downhill=[0, 0, 0, 0])
    for k in range(steps):
        T = t0 * (t1 / t0) ** (k / max(1, steps - 1))
        cand = neighbour(cur, rng)
        val = score_fn(cand)
        if val >= cur_val or rng.uniform() < math.exp((val - cur_val) / T):
            if stats is not None:
                stats["uphill" if val >= cur_val else "downhill"][min(3, 4 * k // max(1, steps))] += 1
            cur, cur_val = cand, val
            accepted += 1
            traced = rescore(cand) if rescore else val
            if traced > best_val:
                best, best_val = cand, traced
    return best, best_val, accepted


def calibrate_temperature(batch_score, seed: int, pairs: int = 32, chi0: float = 0.5, ratio: float = 100.0):
    """Start temperature by Kirkpatrick's rule: a typical downhill move (the
    median |score change| over random (placement, neighbour) pairs, scored by
    the search's own score function) is accepted with probability chi0 at
    the start; the schedule ends at t0 / ratio."""
    rng = SplitMix64(seed)
    ps = [random_placement(rng) for _ in range(pairs)]
    qs = [neighbour(p, rng) for p in ps]
    d = float(np.median(np.abs(np.asarray(batch_score(ps)) - np.asarray(batch_score(qs)))))
    t0 = max(d, 1e-4) / math.log(1.0 / chi0)
    return t0, t0 / ratio


def rank_check(pt: PlacementTracer, net: "SurrogateNet", occ: np.ndarray, n: int = 64, seed: int = 77) -> dict:
    """On n random placements of the search scene: Spearman rank correlation
    of the proxy score (traced matched-phase gain) and of the network's
    prediction with the trial score, and the trial score of the placement
    each of them ranks first against the best trial score."""
    from scipy.stats import spearmanr
    rng = SplitMix64(seed)
    ps = [random_placement(rng) for _ in range(n)]
    trial = np.asarray([pt.trial_score(p) for p in ps])
    proxy = np.asarray([float(pt.gain_map(p).mean()) for p in ps])
    pred = predict_score(net, occ, ps).cpu().numpy() if net is not None else None
    out = {"placements": n, "spearman_proxy_vs_trial": float(spearmanr(proxy, trial)[0]),
           "trial_db_best": float(trial.max()), "trial_db_of_proxy_best": float(trial[int(np.argmax(proxy))]),
           "trial_db_mean": float(trial.mean())}
    if pred is not None:
        out["spearman_prediction_vs_trial"] = float(spearmanr(pred, trial)[0])
        out["trial_db_of_prediction_best"] = float(trial[int(np.argmax(pred))])
    return out


def run_experiment(n_scenes: int = 20, per_scene: int = 1000, sa_steps: int = 2000, search_scene: int = 1000,
                   epochs: int = 20, fib: int = 1 << 18, seed: int = 0, log=print, datasets=None,
                   t0: Optional[float] = None, t1: Optional[float] = None, chi0: float = 0.2,
                   t_ratio: float = 1000.0, net_path: Optional[str] = None, searches=("plain", "surrogate"),
                   judge: str = "trial", rank_n: int = 64) -> dict:
    """Config-4 experiment: traced dataset -> trained surrogate -> placement
    search three ways on an unseen scene: plain annealing (every candidate
    scored by the judge), surrogate-guided annealing (accepted candidates
    scored by the judge), and plain annealing cut to the surrogate search's
    number of judge calls.  judge "trial": PlacementTracer.trial_score (the
    optimised coverage gain a trial reports); "proxy": the mean traced
    matched-phase gain.

    With t0 / t1 None each search calibrates its own schedule with
    calibrate_temperature on its own score function (plain: judged pairs,
    counted in its time and calls; surrogate: predicted pairs)."""
    out: Dict = {}
    t_start = time.perf_counter()
    if datasets:
        P, S_, Y = load_dataset(datasets)
        out["dataset"] = {"files": len(datasets), "seconds": None}
    else:
        P, S_, Y = generate_dataset(n_scenes, per_scene, seed=seed, fib_ap=fib, fib_ms=fib, log=log)
        torch.cuda.synchronize()
        out["dataset"] = {"seconds": time.perf_counter() - t_start}
    X = build_inputs(P, S_)
    Y = Y.reshape(-1, 1, GRID, GRID)
    out["dataset"].update(samples=int(X.shape[0]), scenes=int(len(np.unique(S_))))
    n = X.shape[0]
    n_val = max(64, n // 20)
    Xtr, Ytr, Xva, Yva = X[:-n_val], Y[:-n_val], X[-n_val:], Y[-n_val:]
    net = SurrogateNet(seed=seed)
    tt = time.perf_counter()
    if net_path and os.path.exists(net_path):
        net.load(net_path)
        hist = [float("nan")]
    else:
        hist = train(net, Xtr, Ytr, epochs=epochs, seed=seed, log=log)
        if net_path:
            net.save(net_path)
    torch.cuda.synchronize()
    out["train_seconds"] = time.perf_counter() - tt
    with torch.no_grad():
        pv = torch.cat([net.forward(Xva[i:i + 64]) for i in range(0, n_val, 64)]) * LABEL_SCALE
        mse_val = float(((pv - Yva) ** 2).mean())
        base = float(((Ytr.mean() - Yva) ** 2).mean())
        sv, tv = pv.mean((1, 2, 3)), Yva.mean((1, 2, 3))
        r2_score = 1 - float(((sv - tv) ** 2).sum() / ((tv - tv.mean()) ** 2).sum())
    out.update(train_mse_final_db2=hist[-1] * LABEL_SCALE ** 2, val_mse_db2=mse_val, val_mse_mean_predictor_db2=base,
               val_score_r2=r2_score)
    log(f"validation: map mse {mse_val:.2f} dB^2 (mean predictor {base:.2f}); placement-score R^2 {r2_score:.3f}")
    out.update(search_experiment(net, sa_steps, search_scene, fib, seed, log, chi0, t_ratio, searches, judge, rank_n,
                                 t0, t1))
    return out


def search_experiment(net: "SurrogateNet", sa_steps: int = 2000, search_scene: int = 1000, fib: int = 1 << 18,
                      seed: int = 0, log=print, chi0: float = 0.2, t_ratio: float = 1000.0,
                      searches=("plain", "surrogate"), judge: str = "trial", rank_n: int = 64,
                      t0: Optional[float] = None, t1: Optional[float] = None) -> dict:
    """The searches of run_experiment with a trained network, on the unseen
    scene search_scene: the rank check, plain annealing, surrogate-guided
    annealing and plain annealing cut to the surrogate search's judge calls."""
    out: Dict = {"sa_steps": sa_steps, "search_scene": search_scene, "judge": judge}
    occ = occupancy(search_scene)
    pt = PlacementTracer(search_scene, fib, fib)
    if rank_n:
        out["rank_check"] = rank_check(pt, net, occ, rank_n)
        log(f"rank check on {rank_n} held-out placements: {out['rank_check']}")

    def traced(p):
        if judge == "trial":
            return pt.trial_score(p)
        return float(pt.gain_map(p).mean())

    def timed(fn):
        torch.cuda.synchronize()
        pt.gpu_seconds, pt.n_traces = 0.0, 0
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t, pt.n_traces

    def predicted_batch(ps):
        return predict_score(net, occ, ps).cpu().numpy()

    def search(score_fn, batch_score, steps, rescore=None):
        if t0 is None:
            ta, tb = calibrate_temperature(batch_score, seed + 7, chi0=chi0, ratio=t_ratio)
        else:
            ta, tb = t0, t1
        st: dict = {}
        best, val, acc = anneal(score_fn, steps, seed=seed + 1, t0=ta, t1=tb, rescore=rescore, stats=st)
        return best, val, acc, ta, tb, st

    out["chi0"], out["t_ratio"] = chi0, t_ratio
    if "plain" in searches:
        (b1, v1, a1, ta1, tb1, st1), s1, n1 = timed(lambda: search(traced, lambda ps: [traced(p) for p in ps],
                                                                   sa_steps))
        out["plain"] = {"best_gain_db": v1, "seconds": s1, "traces": n1, "accepted": a1, "t0_db": ta1,
                        "t1_db": tb1, "accepted_by_quarter": st1, "best": b1.__dict__}
        log(f"plain search: {v1:.3f} dB in {s1:.1f} s ({n1} traces, {a1} accepted, t0 {ta1:.3f} dB, {st1})")
    if "surrogate" in searches:
        (b2, v2, a2, ta2, tb2, st2), s2, n2 = timed(lambda: search(lambda p: float(predict_score(net, occ, [p])[0]),
                                                                   predicted_batch, sa_steps, rescore=traced))
        out["surrogate"] = {"best_gain_db": v2, "seconds": s2, "traces": n2, "accepted": a2, "t0_db": ta2,
                            "t1_db": tb2, "accepted_by_quarter": st2, "best": b2.__dict__}
        log(f"surrogate search: {v2:.3f} dB in {s2:.1f} s ({n2} traces, {a2} accepted, t0 {ta2:.3f} dB, {st2})")
    if "plain" in searches and "surrogate" in searches:
        budget = max(1, n2 - 1)
        (b3, v3, a3), s3, n3 = timed(lambda: anneal(traced, budget, seed=seed + 1, t0=ta1, t1=tb1))
        out["plain_same_budget"] = {"best_gain_db": v3, "seconds": s3, "traces": n3, "steps": budget}
        log(f"plain search with the same trace budget: {v3:.3f} dB in {s3:.1f} s")
        out["gain_gap_db"] = v1 - v2
        out["time_ratio"] = s2 / s1
        out["surrogate_gain_over_same_budget_db"] = v2 - v3
    pt.ts.close()
    return out


if __name__ == "__main__":
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="placement surrogate experiment (traced data, training, search)")
    ap.add_argument("--scenes", type=int, default=20)
    ap.add_argument("--per-scene", type=int, default=1000)
    ap.add_argument("--sa-steps", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--fib", type=int, default=1 << 18)
    ap.add_argument("--out", default=None)
    ap.add_argument("--make-dataset", default=None, help="only trace a dataset and save it (npz)")
    ap.add_argument("--first-scene", type=int, default=0)
    ap.add_argument("--datasets", default=None, help="comma-separated npz files to train on")
    ap.add_argument("--net", default=None, help="network weights (.npy): loaded if present, else saved after training")
    ap.add_argument("--chi0", type=float, default=0.2, help="start acceptance of a median downhill move")
    ap.add_argument("--t-ratio", type=float, default=1000.0, help="start / end temperature")
    ap.add_argument("--searches", default="plain,surrogate")
    ap.add_argument("--judge", default="trial", choices=["trial", "proxy"])
    ap.add_argument("--rank-n", type=int, default=64)
    a = ap.parse_args()
    logf = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if a.make_dataset:
        t = time.perf_counter()
        P, S_, Y = generate_dataset(a.scenes, a.per_scene, fib_ap=a.fib, fib_ms=a.fib, log=logf,
                                    first_scene=a.first_scene)
        save_dataset(a.make_dataset, P, S_, Y)
        print(json.dumps({"samples": len(P), "seconds": time.perf_counter() - t}), flush=True)
        sys.exit(0)
    res = run_experiment(a.scenes, a.per_scene, a.sa_steps, epochs=a.epochs, fib=a.fib, log=logf,
                         datasets=a.datasets.split(",") if a.datasets else None, chi0=a.chi0, t_ratio=a.t_ratio,
                         net_path=a.net, searches=tuple(a.searches.split(",")), judge=a.judge, rank_n=a.rank_n)
    txt = json.dumps(res)
    print(txt, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
