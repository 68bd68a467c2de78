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
               splits: int = 1, per_wave: bool = False, split_minor: bool = False,
               three_mult: bool = False, plain_schedule: bool = False) -> torch.Tensor:
    """C[Ra, Rb] = A B^T (or A conj(B)^T) from fragment-packed operands.
    per_wave selects the per-wave kernel also for Ra >= 128; split_minor and
    three_mult the workgroup kernel's split-minor order and three-product
    form (comparisons)."""
    L = _lib.lib()
    out = torch.empty(Ra, Rb, dtype=torch.complex64, device=fa.device)
    ws = None
    if splits > 1:
        ws = torch.empty(max(1, L.ams_cgemm_workspace_bytes(Ra, Rb, splits)), dtype=torch.uint8, device=fa.device)
    _lib.check(L.ams_cgemm_frag(_ptr(fa), _ptr(fb), Ra, Rb, Kd,
                                (1 if conj_b else 0) | (2 if per_wave else 0) | (4 if split_minor else 0)
                                | (8 if three_mult else 0) | (16 if plain_schedule else 0),
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
        # reduction splits of the backward product at M >= 128 (config 3,
        # N = 65,536, 64 AP beams x 4 maps): 16 measured 0.774 ms against
        # 0.804 at 64 (less partial-sum traffic) and 0.920 at 8
        self.bwd_splits = max(1, min(16, (self.N // 32) // 8))

    def nbytes(self) -> int:
        return self.frag_fwd.numel() + self.frag_bwd.numel()


def split_count(Ra: int, Rb: int, Kd: int, cap: int = 64) -> int:
    """Reduction splits of a short-and-wide product on the per-wave kernel
    (a few rows of codes against thousands of columns leave most of the
    GPU's 1024 SIMDs idle otherwise), in waves of 32 x 64 output tiles:
    about 512 waves for short reductions, and for long ones (200 and more
    32-deep blocks) the most splits that keep the waves at or below 1024,
    one per SIMD (just past that, the last waves run as a tail).  Measured:
    30 x 4096 x 1040 forward 49 us unsplit, 19 us at 8 splits; backward
    (30 x 1040 x 4096) 176 us unsplit, 19-21 us at 24-30.  30 x 18900 x
    32036 forward 1.72 ms unsplit, 0.81 ms at 3 splits (1.10 at 4);
    backward 1.12 ms, 0.81 at 2 (0.93 at 3).  30 x 6976 x 65536 backward
    0.98 ms at 4 splits, 0.65 at 9 (0.79 at 12).  30 x 6912 x 6912: 0.35
    ms unsplit, 0.079 at 8 (0.110 at 10)."""
    tiles = ((Ra + 31) // 32) * ((Rb + 63) // 64)
    n_kb = (Kd + 31) // 32
    target = 1024 if n_kb >= 200 else 512
    return max(1, min(cap, n_kb, target // max(1, tiles)))


class Contract(torch.autograd.Function):
    """Y = A . [F ; H_d]  (A: [R, Kc] complex64) with the split-bf16 MFMA GEMM;
    backward gA = gY . [F ; H_d]^H on the same kernels."""

    @staticmethod
    def forward(ctx, A: torch.Tensor, op: ChannelOperand):
        A = A.contiguous()
        R = A.shape[0]
        fa = frag_pack(A, R, op.Kc, op.Kc, 1)
        ctx.op = op
        return cgemm_frag(fa, op.frag_fwd, R, op.N, op.Kc, conj_b=False, splits=split_count(R, op.N, op.Kc))

    @staticmethod
    def backward(ctx, gY: torch.Tensor):
        op = ctx.op
        gY = gY.contiguous()
        R = gY.shape[0]
        fg = frag_pack(gY, R, op.N, op.N, 1)
        splits = split_count(R, op.Kc, op.N) if R < 128 else op.bwd_splits
        gA = cgemm_frag(fg, op.frag_bwd, R, op.Kc, op.N, conj_b=True, splits=splits)
        return gA, None


class PhaseOperand(torch.autograd.Function):
    """The stacked operand A [n_maps * C, K + M] of one metasurface with
    single-bounce chains (go.py:107-163 with max_ms_reflex = 1) from the
    phases, in one kernel (ams_pm_operand); the backward turns gA into the
    phase gradients in two (ams_pm_operand_bwd)."""

    @staticmethod
    def forward(ctx, th_ap: torch.Tensor, th_ms: torch.Tensor, amp: torch.Tensor, ms_amp: torch.Tensor,
                G: torch.Tensor, max_phase: float):
        L = _lib.lib()
        Cn, M = th_ap.shape
        nm, K = th_ms.shape
        dev = th_ap.device
        th_ap, th_ms = th_ap.contiguous(), th_ms.contiguous()
        A = torch.empty(nm * Cn, K + M, dtype=torch.complex64, device=dev)
        WG = torch.empty(Cn, K, dtype=torch.complex64, device=dev)
        U = torch.empty(nm, K, dtype=torch.complex64, device=dev)
        _lib.check(L.ams_pm_operand(_ptr(th_ap), _ptr(amp), _ptr(th_ms), _ptr(ms_amp), _ptr(torch.view_as_real(G)),
                                    Cn, M, K, nm, max_phase, _ptr(A), _ptr(WG), _ptr(U), _stream()),
                   "ams_pm_operand")
        ctx.save_for_backward(th_ap, amp, G, WG, U)
        ctx.dims = (Cn, M, K, nm)
        return A

    @staticmethod
    def backward(ctx, gA: torch.Tensor):
        th_ap, amp, G, WG, U = ctx.saved_tensors
        g_ap, g_ms = _operand_grads(gA, th_ap, amp, G, WG, U, ctx.dims, ctx.needs_input_grad[0])
        return g_ap, g_ms, None, None, None, None


def _operand_grads(gA: torch.Tensor, th_ap, amp, G, WG, U, dims, need_ap: bool):
    """The phase gradients from gA (ams_pm_operand_bwd).  Both land in one
    buffer in FusedAdam's parameter order (AP phases, then metasurface
    phases), so its step needs no concatenation."""
    L = _lib.lib()
    Cn, M, K, nm = dims
    gA = gA.contiguous()
    dev = gA.device
    n_ap = Cn * M if need_ap else 0
    gbuf = torch.empty(n_ap + nm * K, dtype=torch.float32, device=dev)
    g_ap = gbuf[:n_ap].view(Cn, M) if need_ap else None
    g_ms = gbuf[n_ap:].view(nm, K)
    gWG = torch.empty(Cn, K, dtype=torch.complex64, device=dev) if need_ap else None
    _lib.check(L.ams_pm_operand_bwd(_ptr(gA), _ptr(th_ap), _ptr(amp), _ptr(torch.view_as_real(G)), _ptr(WG),
                                    _ptr(U), Cn, M, K, nm, _ptr(g_ms), _ptr(g_ap) if need_ap else None,
                                    _ptr(gWG) if need_ap else None, _stream()),
               "ams_pm_operand_bwd")
    return g_ap, g_ms


class ChainTransfer:
    """The ms-to-ms channels of several metasurfaces as one [K][K] block
    matrix T (K = sum K_b; block (a, b) = ms{a}->ms{b}, block (b, a) its
    transpose, as go.py:147-153 uses it; zero diagonal blocks), packed once in
    the fragment layout.  T is symmetric, so the one packing serves the
    chain products of the forward (P . T) and of the backward
    (gQ . conj(T)^T)."""

    def __init__(self, T: Dict[tuple, torch.Tensor], K_list: Sequence[int], dev):
        off = [0]
        for K in K_list:
            off.append(off[-1] + int(K))
        self.K = off[-1]
        Tf = torch.zeros(self.K, self.K, dtype=torch.complex64, device=dev)
        for (a, b), t in T.items():
            if a < b:
                Tf[off[a]:off[a + 1], off[b]:off[b + 1]] = t
                Tf[off[b]:off[b + 1], off[a]:off[a + 1]] = t.T
        self.frag = frag_pack(Tf, self.K, self.K, self.K, 1)


class ChainOperand(torch.autograd.Function):
    """The stacked operand A [n_maps * C, K + M] of several metasurfaces with
    chains up to `rounds` surfaces long (go.py:107-157): the first round
    (ams_chain_first), each later round's product with the ms-to-ms block
    matrix on the MFMA contraction and its elementwise step
    (ams_chain_step); the backward runs the rounds in reverse
    (ams_chain_bwd_ms, ams_chain_bwd_scale, the conjugate product) and ends
    with the AP phase gradient (ams_chain_bwd_ap).  The metasurfaces'
    elements are concatenated in order."""

    @staticmethod
    def forward(ctx, th_ap: torch.Tensor, amp: torch.Tensor, ms_amp: torch.Tensor, G: torch.Tensor,
                period: torch.Tensor, chain: Optional[ChainTransfer], rounds: int, *th_ms: torch.Tensor):
        L = _lib.lib()
        Cn, M = th_ap.shape
        th = torch.cat([t.detach() for t in th_ms], 1).contiguous()
        nm, K = th.shape
        R, Kc = nm * Cn, K + M
        dev = th_ap.device
        th_ap = th_ap.contiguous()
        A = torch.empty(R, Kc, dtype=torch.complex64, device=dev)
        WG = torch.empty(Cn, K, dtype=torch.complex64, device=dev)
        U = torch.empty(nm, K, dtype=torch.complex64, device=dev)
        _lib.check(L.ams_chain_first(_ptr(th_ap), _ptr(amp), _ptr(th), _ptr(ms_amp), _ptr(period),
                                     _ptr(torch.view_as_real(G)), Cn, M, K, nm, _ptr(A), _ptr(WG), _ptr(U),
                                     _stream()), "ams_chain_first")
        Qs = []
        if chain is not None:
            prev, ld = A, Kc
            for r in range(2, rounds + 1):
                fp = frag_pack(prev, R, K, ld, 1)
                Q = cgemm_frag(fp, chain.frag, R, K, K, conj_b=False, splits=split_count(R, K, K))
                P = torch.empty(R, K, dtype=torch.complex64, device=dev) if r < rounds else None
                _lib.check(L.ams_chain_step(_ptr(Q), _ptr(U), Cn, nm, K, _ptr(A), Kc,
                                            _ptr(P) if P is not None else None, _stream()), "ams_chain_step")
                Qs.append(Q)
                prev, ld = P, K
        ctx.save_for_backward(th_ap, amp, G, WG, U, *Qs)
        ctx.chain = chain
        ctx.dims = (Cn, M, K, nm, [int(t.shape[1]) for t in th_ms])
        return A

    @staticmethod
    def backward(ctx, gA: torch.Tensor):
        L = _lib.lib()
        th_ap, amp, G, WG, U, *Qs = ctx.saved_tensors
        Cn, M, K, nm, Ks = ctx.dims
        R, Kc = nm * Cn, K + M
        gA = gA.contiguous()
        dev = gA.device
        need_ap = ctx.needs_input_grad[0]
        n_ap = Cn * M if need_ap else 0
        gbuf = torch.empty(n_ap + nm * K, dtype=torch.float32, device=dev)
        g_ms = gbuf[n_ap:].view(nm, K)
        gu = torch.empty(nm, K, dtype=torch.complex64, device=dev)
        Gacc = None
        first = True
        for Q in reversed(Qs):
            _lib.check(L.ams_chain_bwd_ms(_ptr(gA), Kc, _ptr(Gacc) if Gacc is not None else None, _ptr(Q), 0,
                                          _ptr(U), Cn, nm, K, _ptr(gu), 0 if first else 1, None, None, _stream()),
                       "ams_chain_bwd_ms")
            first = False
            gQ = torch.empty(R, K, dtype=torch.complex64, device=dev)
            _lib.check(L.ams_chain_bwd_scale(_ptr(gA), Kc, _ptr(Gacc) if Gacc is not None else None, _ptr(U),
                                             Cn, nm, K, _ptr(gQ), _stream()), "ams_chain_bwd_scale")
            fq = frag_pack(gQ, R, K, K, 1)
            Gacc = cgemm_frag(fq, ctx.chain.frag, R, K, K, conj_b=True, splits=split_count(R, K, K))
        gWG = torch.empty(Cn, K, dtype=torch.complex64, device=dev) if need_ap else None
        _lib.check(L.ams_chain_bwd_ms(_ptr(gA), Kc, _ptr(Gacc) if Gacc is not None else None, _ptr(WG), 1,
                                      _ptr(U), Cn, nm, K, _ptr(gu), 0 if first else 1, _ptr(g_ms),
                                      _ptr(gWG) if need_ap else None, _stream()), "ams_chain_bwd_ms")
        g_ap = None
        if need_ap:
            g_ap = gbuf[:n_ap].view(Cn, M)
            _lib.check(L.ams_chain_bwd_ap(_ptr(gA), _ptr(gWG), _ptr(torch.view_as_real(G)), _ptr(th_ap), _ptr(amp),
                                          Cn, M, K, nm, _ptr(g_ap), _stream()), "ams_chain_bwd_ap")
        outs, off = [], 0
        for k in Ks:
            outs.append(g_ms[:, off:off + k] if nm > 1 else g_ms[:, off:off + k].reshape(1, k))
            off += k
        return (g_ap,) + (None,) * 6 + tuple(outs)


def _row_mask(rowmask: Optional[torch.Tensor], groups: int, C: int, dev) -> Optional[torch.Tensor]:
    """The capacity objective's row masks as u8 [groups * C] on the device:
    group g masks its rows rowmask[g * C + c]."""
    if rowmask is None:
        return None
    if rowmask.numel() != groups * C:
        raise ValueError(f"rowmask has {rowmask.numel()} entries; the objective needs one per row "
                         f"({groups} groups x {C} codes)")
    return rowmask.reshape(-1).to(device=dev, dtype=torch.uint8).contiguous()


class PhaseStep(torch.autograd.Function):
    """The whole forward of one metasurface with single-bounce chains (the
    operand from the phases, the contraction, the objective) and its
    backward as one Function.  The operand and gY are written straight into
    the contraction's fragment layout (ams_pm_operand_frag,
    ams_objective_bwd_frag), so neither A nor gY is formed as a matrix.
    Returns the per-map loss and coverage (the latter without gradient)."""

    @staticmethod
    def forward(ctx, th_ap: torch.Tensor, th_ms: torch.Tensor, amp: torch.Tensor, ms_amp: torch.Tensor,
                G: torch.Tensor, max_phase: float, op: ChannelOperand, mode: int, n0: float, w: float,
                rowmask: Optional[torch.Tensor]):
        L = _lib.lib()
        Cn, M = th_ap.shape
        nm, K = th_ms.shape
        R, Kc, N = nm * Cn, op.Kc, op.N
        dev = th_ap.device
        th_ap, th_ms = th_ap.contiguous(), th_ms.contiguous()
        fa = torch.empty(L.ams_frag_bytes(R, Kc), dtype=torch.uint8, device=dev)
        WG = torch.empty(Cn, K, dtype=torch.complex64, device=dev)
        U = torch.empty(nm, K, dtype=torch.complex64, device=dev)
        _lib.check(L.ams_pm_operand_frag(_ptr(th_ap), _ptr(amp), _ptr(th_ms), _ptr(ms_amp),
                                         _ptr(torch.view_as_real(G)), Cn, M, K, nm, max_phase, _ptr(fa), _ptr(WG),
                                         _ptr(U), _stream()),
                   "ams_pm_operand_frag")
        Y = cgemm_frag(fa, op.frag_fwd, R, N, Kc, conj_b=False, splits=split_count(R, N, Kc))
        ws = torch.empty(L.ams_objective_workspace_bytes(nm, N), dtype=torch.uint8, device=dev)
        loss = torch.empty(nm, dtype=torch.float32, device=dev)
        cov = torch.empty(nm, dtype=torch.float32, device=dev)
        rm = _row_mask(rowmask, nm, Cn, dev)
        _lib.check(L.ams_objective_fwd_cov(_ptr(Y), nm, Cn, N, mode, n0, w, _ptr(rm) if rm is not None else None,
                                           _ptr(loss), _ptr(cov), _ptr(ws), _stream()),
                   "ams_objective_fwd_cov")
        ctx.save_for_backward(th_ap, amp, G, WG, U, Y)
        ctx.ws, ctx.rm, ctx.op = ws, rm, op
        ctx.cfg = (Cn, M, K, nm, mode, n0, w)
        ctx.mark_non_differentiable(cov)
        ctx.set_materialize_grads(False)
        return loss, cov

    @staticmethod
    def backward(ctx, gloss: torch.Tensor, gcov: torch.Tensor = None):
        if gloss is None:
            return (None,) * 11
        L = _lib.lib()
        th_ap, amp, G, WG, U, Y = ctx.saved_tensors
        Cn, M, K, nm, mode, n0, w = ctx.cfg
        op = ctx.op
        R, Kc, N = nm * Cn, op.Kc, op.N
        gl = gloss.to(torch.float32).contiguous()
        fg = torch.empty(L.ams_frag_bytes(R, N), dtype=torch.uint8, device=Y.device)
        _lib.check(L.ams_objective_bwd_frag(_ptr(Y), nm, Cn, N, mode, n0, w,
                                            _ptr(ctx.rm) if ctx.rm is not None else None, _ptr(gl), _ptr(ctx.ws),
                                            _ptr(fg), _stream()),
                   "ams_objective_bwd_frag")
        splits = split_count(R, Kc, N) if R < 128 else op.bwd_splits
        gA = cgemm_frag(fg, op.frag_bwd, R, Kc, N, conj_b=True, splits=splits)
        g_ap, g_ms = _operand_grads(gA, th_ap, amp, G, WG, U, (Cn, M, K, nm), ctx.needs_input_grad[0])
        return (g_ap, g_ms) + (None,) * 9


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
    per-group loss [G] and, as a side output without gradient, the per-group
    coverage [G] (mean over Rx of the best code's power in dB).
    mode 0: -mean_r log2(1 + max_c |y|^2 / n0);
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
        cov = torch.empty(groups, dtype=torch.float32, device=Y.device)
        rm = _row_mask(rowmask, groups, Cc, Y.device)
        _lib.check(L.ams_objective_fwd_cov(_ptr(Y), groups, Cc, N, mode, n0, w,
                                           _ptr(rm) if rm is not None else None, _ptr(loss), _ptr(cov), _ptr(ws),
                                           _stream()),
                   "ams_objective_fwd_cov")
        ctx.save_for_backward(Y)
        ctx.ws, ctx.rm, ctx.cfg = ws, rm, (groups, Cc, N, mode, n0, w)
        ctx.mark_non_differentiable(cov)
        ctx.set_materialize_grads(False)
        return loss, cov

    @staticmethod
    def backward(ctx, gloss: torch.Tensor, gcov: torch.Tensor = None):
        if gloss is None:
            return None, None, None, None, None, None
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


def ap_amplitude(code_num: int, M: int) -> torch.Tensor:
    """go.py:59-66: every AP element at amplitude 1/sqrt(M/36); the 4 corners
    of a 6x6 AP are off.  [code_num, M] float32."""
    A = torch.ones(code_num, M, dtype=torch.float32) / math.sqrt(M / 36)
    if M == 36:
        A[:, [0, 5, 30, 35]] = 0
    return A


def start_phases(K: int, M: int, code_num: int = 30, seed: int = 0):
    """The seeded start of the config-1 optimisation: metasurface phases
    [1, K] and AP phases [code_num, M], uniform on [0, 2 pi)."""
    g = torch.Generator().manual_seed(seed)
    th_ms = torch.rand(1, K, generator=g) * 2 * math.pi
    th_ap = torch.rand(code_num, M, generator=g) * 2 * math.pi
    return th_ms, th_ap


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
                 max_ms_reflex: int = 1, fused: bool = True):
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
        self.register_buffer("antenna_A", ap_amplitude(code_num, M).to(dev))
        th = antenna_theta if antenna_theta is not None else torch.rand(code_num, M, generator=g) * 2 * math.pi
        self.antenna_theta = nn.Parameter(th.to(dev, torch.float32).clone(), requires_grad=optimize_codebook)
        self.ms_theta = nn.ParameterList()
        self.ms_A = []
        for i in range(len(ms_list)):
            K = channels["src_to_ms{%d}" % i].shape[1]
            self.ms_A.append(torch.ones(K, dtype=torch.float32, device=dev))
            t0 = ms_theta[i] if ms_theta is not None else torch.zeros(n_maps, K)
            self.ms_theta.append(nn.Parameter(t0.to(dev, torch.float32).reshape(n_maps, K).clone()))
        self.G_list = [channels["src_to_ms{%d}" % i].to(torch.complex64).contiguous() for i in range(len(ms_list))]
        # one metasurface: every chain has length 1, and the operand comes from
        # the fused kernels (PhaseOperand); several: from the chain kernels
        # (ChainOperand); the plain comparison path forms it with torch ops
        self.fused_operand = fused and not plain and len(ms_list) == 1 and max_ms_reflex >= 1
        # the whole step as one Function over fragment-writing kernels (up to
        # 128 AP elements); otherwise the operand, the contraction and the
        # objective as separate Functions
        self.fused_step = self.fused_operand and M <= 128
        # metasurface-to-metasurface channels; ms{b}->ms{a} for b > a is the
        # transpose (not the conjugate) of ms{a}->ms{b}, as go.py:147-153 uses it
        self.T = {}
        for a in range(len(ms_list)):
            for b in range(len(ms_list)):
                if a < b and ("ms{%d}_to_ms{%d}" % (a, b)) in channels:
                    self.T[(a, b)] = channels["ms{%d}_to_ms{%d}" % (a, b)].to(torch.complex64)
                    self.T[(b, a)] = self.T[(a, b)].T
        self.chain_operand = fused and not plain and len(ms_list) > 1 and max_ms_reflex >= 1
        if self.chain_operand:
            self.G_cat = torch.cat(self.G_list, 1).contiguous()
            self.ms_amp_cat = torch.cat(self.ms_A).contiguous()
            self.period = torch.cat([torch.full((int(g.shape[1]),), ms.max_phase_rad, dtype=torch.float32)
                                     for g, ms in zip(self.G_list, ms_list)]).to(dev)
            self.chain = ChainTransfer(self.T, [int(g.shape[1]) for g in self.G_list], dev) \
                if (self.T and max_ms_reflex > 1) else None
        F_list = [channels["ms{%d}_to_dst" % i] for i in range(len(ms_list))]
        self.op = ChannelOperand(F_list, channels["src_to_dst"], keep_plain=plain)
        self.loss_list: List[float] = []
        self.rowmask_fn = None
        self._eff_op, self._eff_key = None, None

    def stacked_operand(self) -> torch.Tensor:
        """A [n_maps * C, Kc] = [ A_0 | A_1 | ... | W ] per map, where A_b
        collects every metasurface chain ending at b (go.py:107-157):
        P1_b = (W G_b) diag(u_b),  Pr_b = sum_{a != b} (P(r-1)_a T_ab) diag(u_b),
        A_b = sum_{r <= max_ms_reflex} Pr_b; so y = A . [F ; H_d]."""
        W = self.antenna_A * torch.exp(-1j * self.antenna_theta)  # go.py:163
        rows = []
        n = len(self.ms_list) if self.max_ms_reflex >= 1 else 0
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
            zeros = [torch.zeros(W.shape[0], g.shape[1], dtype=torch.complex64, device=W.device)
                     for g in self.G_list[n:]]  # max_ms_reflex 0: no chain reaches a metasurface
            rows.append(torch.cat(acc + zeros + [W.to(torch.complex64)], 1))
        return torch.cat(rows, 0)

    @torch.no_grad()
    def postprocess(self):
        """go.py:89-97: AP phases rounded to multiples of pi/2, metasurface
        phases wrapped to [0, max_phase)."""
        self.antenna_theta.copy_(torch.round(self.antenna_theta / (math.pi / 2)) * (math.pi / 2))
        for i, ms in enumerate(self.ms_list):
            self.ms_theta[i].copy_(torch.remainder(self.ms_theta[i], ms.max_phase_rad))
        return self.forward()

    def frozen_operand(self) -> Optional[ChannelOperand]:
        """With one metasurface whose phases are not optimised (one map; the
        reference freezes them for its "MS not optimized" model,
        physical_exp_nni_model.py:136), the metasurface adds a constant to
        the channel: H_eff = H_d + G diag(u) F, formed on the same GEMM once
        per value of the phases (their tensor's version counter), and the
        codebook is optimised against it.  A step captured in a HIP graph
        replays the channel formed at capture: set the phases before
        optimize_model(graph=True).  None otherwise."""
        if not (self.fused_operand and self.n_maps == 1 and not self.ms_theta[0].requires_grad):
            return None
        th = self.ms_theta[0]
        key = (th.data_ptr(), th._version)
        if self._eff_key != key:
            with torch.no_grad():
                u = self.ms_A[0] * torch.exp(1j * torch.remainder(th[0], self.ms_list[0].max_phase_rad))
                M = self.G_list[0].shape[0]
                A = torch.cat([self.G_list[0] * u, torch.eye(M, dtype=torch.complex64, device=th.device)], 1)
                self._eff_op = ChannelOperand([], Contract.apply(A.to(torch.complex64), self.op))
            self._eff_key = key
        return self._eff_op

    def received(self) -> torch.Tensor:
        eff = self.frozen_operand()
        if eff is not None:
            W = self.antenna_A * torch.exp(-1j * self.antenna_theta)  # go.py:163
            return Contract.apply(W.to(torch.complex64), eff)
        if self.fused_operand:
            A = PhaseOperand.apply(self.antenna_theta, self.ms_theta[0], self.antenna_A, self.ms_A[0],
                                   self.G_list[0], self.ms_list[0].max_phase_rad)
        elif self.chain_operand:
            A = ChainOperand.apply(self.antenna_theta, self.antenna_A, self.ms_amp_cat, self.G_cat, self.period,
                                   self.chain, self.max_ms_reflex, *self.ms_theta)
        else:
            A = self.stacked_operand()
        fn = ContractPlain if self.plain else Contract
        return fn.apply(A, self.op)

    def forward(self) -> torch.Tensor:
        rowmask = None
        if self.mode == 1:
            # go.py:220: rows of cap zeroed where torch.rand(C) < 0.1 (CPU generator)
            rowmask = (torch.rand(self.code_num * self.n_maps) < 0.1) if self.rowmask_fn is None \
                else self.rowmask_fn()
        # last_cov: mean over Rx of the best code's power (dB), per map
        if self.fused_step and self.frozen_operand() is None:
            loss, self.last_cov = PhaseStep.apply(self.antenna_theta, self.ms_theta[0], self.antenna_A,
                                                  self.ms_A[0], self.G_list[0], self.ms_list[0].max_phase_rad,
                                                  self.op, self.mode, self.n0, self.w, rowmask)
        else:
            loss, self.last_cov = Objective.apply(self.received(), self.n_maps, self.mode, self.n0, self.w,
                                                  rowmask)
        self.loss = loss.sum() if self.n_maps > 1 else loss.reshape(())
        return self.loss

    @torch.no_grad()
    def coverage_db(self) -> torch.Tensor:
        """Mean over Rx of the best code's received power in dB, per map."""
        Y = self.received().reshape(self.n_maps, self.code_num, -1)
        p = (Y.real ** 2 + Y.imag ** 2).amax(1)
        return (10 * torch.log10(p + 1e-30)).mean(1)


class FusedAdam:
    """torch.optim.Adam semantics (defaults beta=(0.9, 0.999), eps=1e-8) with
    one fused kernel over all parameters (ams_adam_step_dev).  The step count
    and the best-loss tracking live on the device, so a step can be captured
    in a HIP graph."""

    def __init__(self, params, lr: float = 0.3, betas=(0.9, 0.999), eps: float = 1e-8):
        self.params = [p for p in params if p.requires_grad]
        self.lr, self.b1, self.b2, self.eps = lr, betas[0], betas[1], eps
        n = sum(p.numel() for p in self.params)
        dev = self.params[0].device
        self.flat = torch.cat([p.detach().reshape(-1) for p in self.params]).contiguous()
        self.m = torch.zeros(n, dtype=torch.float32, device=dev)
        self.v = torch.zeros(n, dtype=torch.float32, device=dev)
        # steps taken, the step's bias corrections, best-loss tracking
        self.state = torch.zeros(_lib.AMS_ADAM_STATE_BYTES // 4, dtype=torch.int32, device=dev)
        off = 0
        for p in self.params:  # parameters become views of one flat buffer
            k = p.numel()
            p.data = self.flat[off:off + k].view_as(p)
            off += k

    @property
    def steps_taken(self) -> int:
        return int(self.state[0].item())

    def zero_grad(self):
        for p in self.params:
            p.grad = None

    def _flat_grad(self) -> torch.Tensor:
        """The gradients as one float32 vector in self.flat's order: a view
        when they already lie back to back in one buffer (PhaseOperand's
        backward writes them so), else their concatenation."""
        grads = [p.grad for p in self.params]
        if all(gr is not None and gr.dtype == torch.float32 and gr.is_contiguous() for gr in grads):
            st = grads[0].untyped_storage()
            off = grads[0].storage_offset()
            ok = True
            for p, gr in zip(self.params, grads):
                ok = ok and gr.untyped_storage().data_ptr() == st.data_ptr() and gr.storage_offset() == off
                off += p.numel()
            if ok and off * 4 <= st.nbytes():
                return grads[0].new_empty(0).set_(st, grads[0].storage_offset(), (self.flat.numel(),), (1,))
        g = torch.cat([(gr if gr is not None else torch.zeros_like(p)).reshape(-1)
                       for p, gr in zip(self.params, grads)])
        return g.to(torch.float32).contiguous()

    @torch.no_grad()
    def step(self, loss: Optional[torch.Tensor] = None, best: Optional[torch.Tensor] = None,
             hist: Optional[torch.Tensor] = None):
        """One update.  With loss (the device scalar of the parameters before
        this update), parameters whose loss is the lowest so far are copied
        to best (a flat buffer like self.flat) and the loss is written to
        hist[step] (both optional)."""
        g = self._flat_grad()
        if loss is not None:
            loss = loss.detach().to(torch.float32)
        if best is not None and (best.numel() != self.flat.numel() or best.dtype != torch.float32
                                 or not best.is_contiguous()):
            raise ValueError("FusedAdam.step: best must be a contiguous float32 buffer of the parameters' size")
        _lib.check(_lib.lib().ams_adam_step_dev(
            _ptr(self.flat), _ptr(g), _ptr(self.m), _ptr(self.v), self.flat.numel(), self.lr, self.b1, self.b2,
            self.eps, _ptr(self.state), _ptr(loss) if loss is not None else None,
            _ptr(best) if best is not None else None, _ptr(hist) if hist is not None else None,
            hist.numel() if hist is not None else 0, _stream()), "ams_adam_step_dev")


def optimize_model(model: PhaseMapModel, lr: float = 0.3, step: int = 200, fused: bool = True,
                   record: bool = True, callback=None, graph: bool = False, graph_dot: Optional[str] = None):
    """Adam over the model parameters keeping the best-loss parameters
    (go.py:296-336).  The best loss is tracked on the device (no per-step
    host sync; inside the fused Adam when fused); the loss history is read
    back once at the end.

    graph=True captures one whole step (forward, the hand-written backward
    and the fused Adam with its tracking) in a HIP graph and replays it:
    the step's small launches become one graph launch.  It needs the fused
    Adam.  The capacity objective draws a row mask per step from torch's CPU
    generator (go.py:220): with graph=True all `step` masks are drawn first,
    in the same order and so with the same values, kept on the device, and
    each step picks its own by the Adam step count on the device; a
    rowmask_fn of the caller cannot be replayed and is refused.  graph_dot
    names a file for the captured graph's nodes (Graphviz DOT, from the
    runtime's graph dump)."""
    params = [p for p in model.parameters() if p.requires_grad]
    if graph:
        if not fused:
            raise ValueError("optimize_model: graph=True needs the fused Adam")
        if model.mode != 0 and model.rowmask_fn is not None:
            raise ValueError("optimize_model: graph=True cannot replay a caller's rowmask_fn; leave it unset and "
                             "the row masks are drawn as go.py:220 draws them")
    if not fused:
        return _optimize_torch_adam(model, params, lr, step, record, callback)
    opt = FusedAdam(params, lr=lr)
    dev = opt.flat.device
    best = opt.flat.clone()
    hist = torch.zeros(max(step, 1), dtype=torch.float32, device=dev) if record else None
    seed = torch.ones((), dtype=torch.float32, device=dev)  # d loss / d loss, made once

    def body():
        loss = model()
        loss.backward(seed)
        opt.step(loss=loss, best=best, hist=hist)
        opt.zero_grad()

    if graph and model.mode != 0 and step > 0:
        # the eager loop's masks, drawn up front: row i for step i, chosen on
        # the device by the step count (the fused Adam's state[0])
        C = model.code_num * model.n_maps
        masks = torch.stack([torch.rand(C) < 0.1 for _ in range(step)]).to(dev)
        cur_mask = torch.empty(C, dtype=torch.bool, device=dev)

        def mask_of_step():
            cur_mask.copy_(masks.index_select(0, opt.state[:1].long()).reshape(-1))
            return cur_mask

        model.rowmask_fn = mask_of_step
        try:
            _replay_steps(model, body, step, callback, graph_dot)
        finally:
            model.rowmask_fn = None
    elif graph:
        _replay_steps(model, body, step, callback, graph_dot)
    else:
        for it in range(step):
            body()
            if callback is not None:
                callback(it, model)
    with torch.no_grad():
        opt.flat.copy_(best)
    if record and step > 0:
        model.loss_list.extend(hist[:step].cpu().tolist())
    return model


def _optimize_torch_adam(model: PhaseMapModel, params, lr: float, step: int, record: bool, callback):
    """optimize_model with torch.optim.Adam; best-loss tracking by torch ops."""
    opt = torch.optim.Adam(params, lr=lr)
    best = torch.tensor(float("inf"), device=params[0].device)
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


_CAPTURE_STREAMS: Dict[int, torch.cuda.Stream] = {}


def _capture_stream(dev) -> torch.cuda.Stream:
    """One side stream per device for warm-up and capture, made once: a new
    stream costs the libraries a new workspace each time."""
    i = torch.device(dev).index
    i = torch.cuda.current_device() if i is None else i
    if i not in _CAPTURE_STREAMS:
        _CAPTURE_STREAMS[i] = torch.cuda.Stream(device=i)
    return _CAPTURE_STREAMS[i]


def _replay_steps(model: PhaseMapModel, body, step: int, callback, graph_dot: Optional[str] = None):
    """Run body() `step` times with the step captured once and replayed.  The
    first step runs eagerly on a side stream (capture needs whatever the step
    initialises lazily done first); it is a real step of the run.  The second
    is captured on that stream and every step from there on is a replay, so
    the trajectory is the eager loop's: the same kernels in the same order,
    all state (step count, best loss, history) on the device."""
    dev = model.antenna_theta.device
    cur = torch.cuda.current_stream(dev)
    side = _capture_stream(dev)
    n_eager = min(1, step)
    for it in range(n_eager):
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            body()
        cur.wait_stream(side)
        # drop the step's autograd graph, so the parameters' gradient
        # accumulators are made afresh on the capture stream
        model.loss = None
        if callback is not None:
            callback(it, model)
    if step <= n_eager:
        return
    g = torch.cuda.CUDAGraph()
    if graph_dot:
        g.enable_debug_mode()
    side.wait_stream(cur)
    with torch.cuda.stream(side):
        g.capture_begin()
        try:
            body()
        finally:
            g.capture_end()
    if graph_dot:
        g.debug_dump(graph_dot)
    for it in range(n_eager, step):
        g.replay()
        if callback is not None:
            callback(it, model)


def channels_from_traces(H_d: torch.Tensor, G: torch.Tensor, F: torch.Tensor) -> Dict[str, torch.Tensor]:
    """The channel dictionary of build_channels for one metasurface
    (go.py:1172-1224 key names)."""
    return {"src_to_dst": H_d, "src_to_ms{0}": G, "ms{0}_to_dst": F}


def time_to_target(channels: Dict[str, torch.Tensor], ms: Metasurface, steps: int = 500, lr: float = 0.3,
                   code_num: int = 30, seed: int = 0, tol_db: float = 0.1, graph: bool = True) -> dict:
    """Optimise the phase map twice from the same start: plain FP32 (vector
    ALU contraction, torch.optim.Adam) and the split-bf16 MFMA path with the
    fused Adam (its step captured in a HIP graph when graph is set).  Returns
    the plain optimum's coverage, and the wall seconds the MFMA path needs to
    come within tol_db of it (step events on the stream, counted from before
    the warm-up and capture)."""
    K = channels["src_to_ms{0}"].shape[1]
    M = channels["src_to_dst"].shape[0]
    th_ms, th_ap = start_phases(K, M, code_num, seed)
    out = {}
    for name, plain in (("plain_fp32", True), ("mfma_split_bf16", False)):
        # untimed: 3 steps of the same optimisation on a throwaway model, so the
        # process's one-time set-up (code loading, streams, allocator pools)
        # stays out of the timed run; the timed run does its own warm-up steps
        # and capture
        warm = PhaseMapModel(name, channels, [ms], code_num=code_num, ms_theta=[th_ms], antenna_theta=th_ap,
                             plain=plain)
        optimize_model(warm, lr=lr, step=3, fused=not plain, graph=graph and not plain)
        del warm
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
        optimize_model(model, lr=lr, step=steps, fused=not plain, callback=cb, graph=graph and not plain)
        torch.cuda.synchronize()
        cov = torch.stack(covs).reshape(-1).cpu()
        t = [e0.elapsed_time(e) / 1e3 for e in evs]
        best_loss_step = int(torch.tensor(model.loss_list).argmin())
        out[name] = {"cov_db": cov, "t": t, "best_cov_db": float(cov[best_loss_step]),
                     "final_loss": float(min(model.loss_list)), "seconds": t[-1],
                     "ms_per_step": (t[-1] - t[0]) / max(1, len(t) - 1) * 1e3}
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
        "mfma_ms_per_step": out["mfma_split_bf16"]["ms_per_step"],
        "plain_ms_per_step": out["plain_fp32"]["ms_per_step"],
        "mfma_graph": bool(graph),
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
