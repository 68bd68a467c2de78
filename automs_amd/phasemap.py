"""Phase-map optimiser on MI355X: the contraction H = H_d + G diag(u) F, the
AP codebook product and the coverage objective, with hand-written forward
and backward kernels (libams.so) behind torch.autograd.Function.

Mirrors go.py's Metasurface / Model / optimize_model (go.py:17-336) for
metasurface chains of length one (max_ms_reflex <= 1), which is what
physical_exp_nni_model.py runs (:113-128).

Reassociation: y = W (H_d + G diag(u) F) = [ (W G) diag(u) | W ] . [ F ; H_d ],
so one complex GEMM against the stacked channel operand [F ; H_d] yields the
received signals of every code; its large operand is packed once into the
split-bf16 MFMA layout (two copies: reduction over the metasurface index for
the forward, over Rx for the backward).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import torch
import torch.nn as nn

from . import _lib

N0_90DBM = 10 ** (-90 / 10)  # go.py:177-178, noise floor -90 dB (mW)


def _stream(dev=None) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def frag_pack(X: torch.Tensor, R: int, Kd: int, sr: int, sk: int) -> torch.Tensor:
    """Pack complex64 X (element (r, k) at offset r*sr + k*sk) into the
    split-bf16 fragment layout; returns a uint8 device buffer."""
    L = _lib.lib()
    nbytes = L.ams_frag_bytes(R, Kd)
    out = torch.empty(nbytes, dtype=torch.uint8, device=X.device)
    xr = torch.view_as_real(X) if X.is_complex() else X
    _lib.check(L.ams_frag_pack(_ptr(xr), R, Kd, sr, sk, _ptr(out), _stream()),
               "ams_frag_pack")
    return out


def cgemm_frag(fa: torch.Tensor, fb: torch.Tensor, Ra: int, Rb: int, Kd: int, conj_b: bool,
               splits: int = 1, per_wave: bool = False) -> torch.Tensor:
    """C[Ra, Rb] = A B^T (or A conj(B)^T) from fragment-packed operands.
    per_wave selects the per-wave kernel also for Ra >= 128 (comparison)."""
    L = _lib.lib()
    out = torch.empty(Ra, Rb, dtype=torch.complex64, device=fa.device)
    ws = None
    if splits > 1:
        ws = torch.empty(max(1, L.ams_cgemm_workspace_bytes(Ra, Rb, splits)), dtype=torch.uint8, device=fa.device)
    _lib.check(L.ams_cgemm_frag(_ptr(fa), _ptr(fb), Ra, Rb, Kd, (1 if conj_b else 0) | (2 if per_wave else 0),
                                _ptr(out), Rb, splits,
                                _ptr(ws) if ws is not None else None, _stream()), "ams_cgemm_frag")
    return out


def cgemm_f32(A: torch.Tensor, B: torch.Tensor, conj_b: bool) -> torch.Tensor:
    """Plain FP32 C = A B^T (complex64, row-major, contiguous)."""
    L = _lib.lib()
    A = A.contiguous()
    B = B.contiguous()
    Ra, Kd = A.shape
    Rb = B.shape[0]
    out = torch.empty(Ra, Rb, dtype=torch.complex64, device=A.device)
    _lib.check(L.ams_cgemm_f32(_ptr(A), Kd, _ptr(B), Kd, Ra, Rb, Kd, 1 if conj_b else 0, _ptr(out), Rb, _stream()),
               "ams_cgemm_f32")
    return out


class ChannelOperand:
    """The stacked channel operand [F_0; ...; F_{m-1}; H_d] of shape [Kc, N]
    packed once in both fragment layouts (kept resident in HBM)."""

    def __init__(self, F_list: Sequence[torch.Tensor], H_d: torch.Tensor, keep_plain: bool = False):
        parts = [f.to(torch.complex64) for f in F_list] + [H_d.to(torch.complex64)]
        Fcat = torch.cat(parts, 0).contiguous()
        self.Kc, self.N = Fcat.shape
        self.K_list = [int(f.shape[0]) for f in F_list]
        self.M = int(H_d.shape[0])
        self.device = Fcat.device
        # forward: rows = Rx, reduction over the stacked index
        self.frag_fwd = frag_pack(Fcat, self.N, self.Kc, 1, self.N)
        # backward: rows = stacked index, reduction over Rx
        self.frag_bwd = frag_pack(Fcat, self.Kc, self.N, self.N, 1)
        self.plain = Fcat if keep_plain else None
        self.bwd_splits = max(1, min(64, (self.N // 32) // 8))

    def nbytes(self) -> int:
        return self.frag_fwd.numel() + self.frag_bwd.numel()


class Contract(torch.autograd.Function):
    """Y = A . [F ; H_d]  (A: [R, Kc] complex64) with the split-bf16 MFMA GEMM;
    backward gA = gY . [F ; H_d]^H on the same kernels."""

    @staticmethod
    def forward(ctx, A: torch.Tensor, op: ChannelOperand):
        A = A.contiguous()
        R = A.shape[0]
        fa = frag_pack(A, R, op.Kc, op.Kc, 1)
        ctx.op = op
        return cgemm_frag(fa, op.frag_fwd, R, op.N, op.Kc, conj_b=False)

    @staticmethod
    def backward(ctx, gY: torch.Tensor):
        op = ctx.op
        gY = gY.contiguous()
        R = gY.shape[0]
        fg = frag_pack(gY, R, op.N, op.N, 1)
        gA = cgemm_frag(fg, op.frag_bwd, R, op.Kc, op.N, conj_b=True, splits=op.bwd_splits)
        return gA, None


class ContractPlain(torch.autograd.Function):
    """The same product in plain FP32 (vector ALU kernel), for comparison."""

    @staticmethod
    def forward(ctx, A: torch.Tensor, op: ChannelOperand):
        ctx.op = op
        return cgemm_f32(A, op.plain.t().contiguous(), conj_b=False)

    @staticmethod
    def backward(ctx, gY: torch.Tensor):
        op = ctx.op
        return cgemm_f32(gY, op.plain, conj_b=True), None


class Objective(torch.autograd.Function):
    """Coverage objective over Y [G*C, N] (G groups of C codes): returns the
    per-group loss [G].  mode 0: -mean_r log2(1 + max_c |y|^2 / n0);
    mode 1: the reference capacity loss (go.py:214-225)."""

    @staticmethod
    def forward(ctx, Y: torch.Tensor, groups: int, mode: int, n0: float, w: float,
                rowmask: Optional[torch.Tensor]):
        L = _lib.lib()
        Y = Y.contiguous()
        GC, N = Y.shape
        Cc = GC // groups
        ws = torch.empty(L.ams_objective_workspace_bytes(groups, N), dtype=torch.uint8, device=Y.device)
        loss = torch.empty(groups, dtype=torch.float32, device=Y.device)
        rm = rowmask.to(device=Y.device, dtype=torch.uint8).contiguous() if rowmask is not None else None
        _lib.check(L.ams_objective_fwd(_ptr(Y), groups, Cc, N, mode, n0, w,
                                       _ptr(rm) if rm is not None else None, _ptr(loss), _ptr(ws), _stream()),
                   "ams_objective_fwd")
        ctx.save_for_backward(Y)
        ctx.ws, ctx.rm, ctx.cfg = ws, rm, (groups, Cc, N, mode, n0, w)
        return loss

    @staticmethod
    def backward(ctx, gloss: torch.Tensor):
        L = _lib.lib()
        (Y,) = ctx.saved_tensors
        groups, Cc, N, mode, n0, w = ctx.cfg
        gY = torch.empty_like(Y)
        gl = gloss.to(torch.float32).contiguous()
        _lib.check(L.ams_objective_bwd(_ptr(Y), groups, Cc, N, mode, n0, w,
                                       _ptr(ctx.rm) if ctx.rm is not None else None, _ptr(gl), _ptr(ctx.ws),
                                       _ptr(gY), _stream()),
                   "ams_objective_bwd")
        return gY, None, None, None, None, None


@dataclass
class Metasurface:
    """go.py:17-24; max_phase in degrees."""
    pos: torch.Tensor
    size: tuple
    send_dir: torch.Tensor
    recv_dir: torch.Tensor
    blackhole: list
    max_phase: float = 360.0

    @property
    def max_phase_rad(self) -> float:
        return self.max_phase / 180 * math.pi


class PhaseMapModel(nn.Module):
    """Codebook + metasurface phase model (go.py:27-245, chains of length <= 1).

    channels: {"src_to_dst": H_d [M, N], "src_to_ms{i}": G_i [M, K_i],
               "ms{i}_to_dst": F_i [K_i, N]} complex on the GPU.
    n_maps > 1 optimises that many metasurface phase maps at once (rows of
    the GEMM stacked as (map, code)); the loss is their sum.
    """

    def __init__(self, name: str, channels: Dict[str, torch.Tensor], ms_list: List[Metasurface],
                 code_num: int = 30, w: float = 1.0, objective: str = "logsnr", n0: float = N0_90DBM,
                 n_maps: int = 1, optimize_codebook: bool = True, antenna_theta: Optional[torch.Tensor] = None,
                 ms_theta: Optional[List[torch.Tensor]] = None, plain: bool = False, seed: int = 0,
                 max_ms_reflex: int = 1):
        super().__init__()
        self.max_ms_reflex = max_ms_reflex
        dev = channels["src_to_dst"].device
        self.name = name
        self.ms_list = ms_list
        self.w = w
        self.n0 = n0
        self.mode = {"logsnr": 0, "capacity": 1}[objective]
        self.n_maps = n_maps
        self.plain = plain
        M = channels["src_to_dst"].shape[0]
        self.code_num = code_num
        g = torch.Generator(device="cpu").manual_seed(seed)
        # go.py:59-66: amplitude 1/sqrt(M/36); the 4 corners of a 6x6 AP are off
        A = torch.ones(code_num, M, dtype=torch.float32) / math.sqrt(M / 36)
        if M == 36:
            A[:, [0, 5, 30, 35]] = 0
        self.register_buffer("antenna_A", A.to(dev))
        th = antenna_theta if antenna_theta is not None else torch.rand(code_num, M, generator=g) * 2 * math.pi
        self.antenna_theta = nn.Parameter(th.to(dev, torch.float32).clone(), requires_grad=optimize_codebook)
        self.ms_theta = nn.ParameterList()
        self.ms_A = []
        for i in range(len(ms_list)):
            K = channels["src_to_ms{%d}" % i].shape[1]
            self.ms_A.append(torch.ones(K, dtype=torch.float32, device=dev))
            t0 = ms_theta[i] if ms_theta is not None else torch.zeros(n_maps, K)
            self.ms_theta.append(nn.Parameter(t0.to(dev, torch.float32).reshape(n_maps, K).clone()))
        self.G_list = [channels["src_to_ms{%d}" % i].to(torch.complex64) for i in range(len(ms_list))]
        # metasurface-to-metasurface channels; ms{b}->ms{a} for b > a is the
        # transpose (not the conjugate) of ms{a}->ms{b}, as go.py:147-153 uses it
        self.T = {}
        for a in range(len(ms_list)):
            for b in range(len(ms_list)):
                if a < b and ("ms{%d}_to_ms{%d}" % (a, b)) in channels:
                    self.T[(a, b)] = channels["ms{%d}_to_ms{%d}" % (a, b)].to(torch.complex64)
                    self.T[(b, a)] = self.T[(a, b)].T
        F_list = [channels["ms{%d}_to_dst" % i] for i in range(len(ms_list))]
        self.op = ChannelOperand(F_list, channels["src_to_dst"], keep_plain=plain)
        self.loss_list: List[float] = []
        self.rowmask_fn = None

    def stacked_operand(self) -> torch.Tensor:
        """A [n_maps * C, Kc] = [ A_0 | A_1 | ... | W ] per map, where A_b
        collects every metasurface chain ending at b (go.py:107-157):
        P1_b = (W G_b) diag(u_b),  Pr_b = sum_{a != b} (P(r-1)_a T_ab) diag(u_b),
        A_b = sum_{r <= max_ms_reflex} Pr_b; so y = A . [F ; H_d]."""
        W = self.antenna_A * torch.exp(-1j * self.antenna_theta)  # go.py:163
        rows = []
        n = len(self.ms_list)
        for m in range(self.n_maps):
            u = []
            for i, ms in enumerate(self.ms_list):
                phase = torch.remainder(self.ms_theta[i][m], ms.max_phase_rad)  # go.py:131
                u.append(self.ms_A[i] * torch.exp(1j * phase))
            P = [(W @ self.G_list[i]) * u[i] for i in range(n)]
            acc = list(P)
            for _ in range(2, self.max_ms_reflex + 1):
                P = [sum(P[a] @ self.T[(a, b)] for a in range(n) if a != b and (a, b) in self.T) * u[b]
                     if any(a != b and (a, b) in self.T for a in range(n)) else torch.zeros_like(P[b])
                     for b in range(n)]
                acc = [acc[b] + P[b] for b in range(n)]
            rows.append(torch.cat(acc + [W.to(torch.complex64)], 1))
        return torch.cat(rows, 0)

    @torch.no_grad()
    def postprocess(self):
        """go.py:89-97: AP phases rounded to multiples of pi/2, metasurface
        phases wrapped to [0, max_phase)."""
        self.antenna_theta.copy_(torch.round(self.antenna_theta / (math.pi / 2)) * (math.pi / 2))
        for i, ms in enumerate(self.ms_list):
            self.ms_theta[i].copy_(torch.remainder(self.ms_theta[i], ms.max_phase_rad))
        return self.forward()

    def received(self) -> torch.Tensor:
        A = self.stacked_operand()
        fn = ContractPlain if self.plain else Contract
        return fn.apply(A, self.op)

    def forward(self) -> torch.Tensor:
        Y = self.received()
        with torch.no_grad():  # coverage side output: mean over Rx of best-code power (dB)
            p = (Y.real ** 2 + Y.imag ** 2).reshape(self.n_maps, self.code_num, -1).amax(1)
            self.last_cov = (10 * torch.log10(p + 1e-30)).mean(1)
        rowmask = None
        if self.mode == 1:
            # go.py:220: rows of cap zeroed where torch.rand(C) < 0.1 (CPU generator)
            rowmask = (torch.rand(self.code_num * self.n_maps) < 0.1) if self.rowmask_fn is None \
                else self.rowmask_fn()
        loss = Objective.apply(Y, self.n_maps, self.mode, self.n0, self.w, rowmask)
        self.loss = loss.sum()
        return self.loss

    @torch.no_grad()
    def coverage_db(self) -> torch.Tensor:
        """Mean over Rx of the best code's received power in dB, per map."""
        Y = self.received().reshape(self.n_maps, self.code_num, -1)
        p = (Y.real ** 2 + Y.imag ** 2).amax(1)
        return (10 * torch.log10(p + 1e-30)).mean(1)


class FusedAdam:
    """torch.optim.Adam semantics (defaults beta=(0.9, 0.999), eps=1e-8) with
    one fused kernel over all parameters (ams_adam_step)."""

    def __init__(self, params, lr: float = 0.3, betas=(0.9, 0.999), eps: float = 1e-8):
        self.params = [p for p in params if p.requires_grad]
        self.lr, self.b1, self.b2, self.eps = lr, betas[0], betas[1], eps
        n = sum(p.numel() for p in self.params)
        dev = self.params[0].device
        self.flat = torch.cat([p.detach().reshape(-1) for p in self.params]).contiguous()
        self.m = torch.zeros(n, dtype=torch.float32, device=dev)
        self.v = torch.zeros(n, dtype=torch.float32, device=dev)
        off = 0
        for p in self.params:  # parameters become views of one flat buffer
            k = p.numel()
            p.data = self.flat[off:off + k].view_as(p)
            off += k
        self.t = 0

    def zero_grad(self):
        for p in self.params:
            p.grad = None

    @torch.no_grad()
    def step(self):
        self.t += 1
        g = torch.cat([(p.grad if p.grad is not None else torch.zeros_like(p)).reshape(-1) for p in self.params])
        g = g.to(torch.float32).contiguous()
        _lib.check(_lib.lib().ams_adam_step(_ptr(self.flat), _ptr(g), _ptr(self.m), _ptr(self.v), self.flat.numel(),
                                            self.lr, self.b1, self.b2, self.eps, self.t, _stream()),
                   "ams_adam_step")


def optimize_model(model: PhaseMapModel, lr: float = 0.3, step: int = 200, fused: bool = True,
                   record: bool = True, callback=None):
    """Adam over the model parameters keeping the best-loss parameters
    (go.py:296-336).  The best loss is tracked on the device (no per-step
    host sync); the loss history is read back once at the end."""
    params = [p for p in model.parameters() if p.requires_grad]
    opt = FusedAdam(params, lr=lr) if fused else torch.optim.Adam(params, lr=lr)
    best = torch.tensor(float("inf"), device=model.antenna_theta.device)
    best_state = [p.detach().clone() for p in params]
    hist = []
    for it in range(step):
        loss = model()
        with torch.no_grad():
            better = loss.detach() < best
            best = torch.where(better, loss.detach(), best)
            for b, p in zip(best_state, params):
                b.copy_(torch.where(better, p.detach(), b))
        if record:
            hist.append(loss.detach())
        loss.backward()
        opt.step()
        opt.zero_grad()
        if callback is not None:
            callback(it, model)
    with torch.no_grad():
        for b, p in zip(best_state, params):
            p.copy_(b)
    if record and hist:
        model.loss_list.extend(torch.stack(hist).cpu().tolist())
    return model


def channels_from_traces(H_d: torch.Tensor, G: torch.Tensor, F: torch.Tensor) -> Dict[str, torch.Tensor]:
    """The channel dictionary of build_channels for one metasurface
    (go.py:1172-1224 key names)."""
    return {"src_to_dst": H_d, "src_to_ms{0}": G, "ms{0}_to_dst": F}


def time_to_target(channels: Dict[str, torch.Tensor], ms: Metasurface, steps: int = 500, lr: float = 0.3,
                   code_num: int = 30, seed: int = 0, tol_db: float = 0.1) -> dict:
    """Optimise the phase map twice from the same start: plain FP32 (vector
    ALU contraction, torch.optim.Adam) and the split-bf16 MFMA path with the
    fused Adam.  Returns the plain optimum's coverage, and the wall seconds the
    MFMA path needs to come within tol_db of it (step events on the stream)."""
    K = channels["src_to_ms{0}"].shape[1]
    M = channels["src_to_dst"].shape[0]
    g = torch.Generator().manual_seed(seed)
    th_ms = torch.rand(1, K, generator=g) * 2 * math.pi
    th_ap = torch.rand(code_num, M, generator=g) * 2 * math.pi
    out = {}
    for name, plain in (("plain_fp32", True), ("mfma_split_bf16", False)):
        model = PhaseMapModel(name, channels, [ms], code_num=code_num, ms_theta=[th_ms], antenna_theta=th_ap,
                              plain=plain)
        covs, evs = [], []
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e0.record()

        def cb(it, m):
            covs.append(m.last_cov.detach().clone())
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            evs.append(e)
        optimize_model(model, lr=lr, step=steps, fused=not plain, callback=cb)
        torch.cuda.synchronize()
        cov = torch.stack(covs).reshape(-1).cpu()
        t = [e0.elapsed_time(e) / 1e3 for e in evs]
        best_loss_step = int(torch.tensor(model.loss_list).argmin())
        out[name] = {"cov_db": cov, "t": t, "best_cov_db": float(cov[best_loss_step]),
                     "final_loss": float(min(model.loss_list)), "seconds": t[-1]}
    target = out["plain_fp32"]["best_cov_db"]
    cov = out["mfma_split_bf16"]["cov_db"]
    hit = (cov >= target - tol_db).nonzero()
    k = int(hit[0]) if len(hit) else None
    return {
        "plain_optimum_cov_db": target,
        "plain_seconds_500": out["plain_fp32"]["seconds"],
        "mfma_seconds_500": out["mfma_split_bf16"]["seconds"],
        "mfma_best_cov_db": out["mfma_split_bf16"]["best_cov_db"],
        "mfma_steps_to_target": None if k is None else k + 1,
        "seconds_to_target": None if k is None else out["mfma_split_bf16"]["t"][k],
        "start_cov_db": float(cov[0]),
        "final_loss_plain": out["plain_fp32"]["final_loss"],
        "final_loss_mfma": out["mfma_split_bf16"]["final_loss"],
    }


def dft_beams(n0: int, n1: int, n_beams: int) -> torch.Tensor:
    """Phases [n_beams, n0*n1] of steering vectors of an n0 x n1 half-wave
    array towards n_beams directions on a regular (u, v) grid in the visible
    region (u^2 + v^2 < 1); the codebook of the batched optimiser."""
    side = int(math.ceil(math.sqrt(n_beams)))
    us = torch.linspace(-0.8, 0.8, side, dtype=torch.float64)
    dirs = [(u, v) for u in us for v in us][:n_beams]
    i = torch.arange(n0, dtype=torch.float64).repeat_interleave(n1)
    j = torch.arange(n1, dtype=torch.float64).repeat(n0)
    ph = torch.stack([math.pi * (u * i + v * j) for u, v in dirs])
    return torch.remainder(ph, 2 * math.pi).to(torch.float32)


def batched_model(channels: Dict[str, torch.Tensor], ms: Metasurface, ap_shape=(8, 8), n_beams: int = 64,
                  n_maps: int = 4, seed: int = 0, plain: bool = False) -> PhaseMapModel:
    """n_maps metasurface phase maps x n_beams fixed AP beams in one GEMM:
    the rows of A are (map, beam); only the phase maps are optimised."""
    K = channels["src_to_ms{0}"].shape[1]
    g = torch.Generator().manual_seed(seed)
    th_ms = torch.rand(n_maps, K, generator=g) * 2 * math.pi
    th_ap = dft_beams(ap_shape[0], ap_shape[1], n_beams)
    return PhaseMapModel("batched", channels, [ms], code_num=n_beams, n_maps=n_maps, ms_theta=[th_ms],
                         antenna_theta=th_ap, optimize_codebook=False, plain=plain)
