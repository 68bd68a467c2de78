"""Torch-CPU (float64 / complex128) restatement of the reference's phase-map
model, used as the oracle of the contraction, objective and gradients.

TEST INFRASTRUCTURE ONLY (tests/, bench.py cpu_baseline).

Follows go.py:107-164 (get_channel for chains of length one, AP codebook
product) and go.py:166-225 (capacity loss), and adds the mean-log-SNR
objective of the synthetic workloads.  Gradients come from torch autograd on
the CPU; torch is an installed library, not the reference.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import torch

EPS32 = float(torch.finfo(torch.float32).eps)


def received(antenna_A: torch.Tensor, antenna_theta: torch.Tensor, H_d: torch.Tensor,
             G_list: Sequence[torch.Tensor], F_list: Sequence[torch.Tensor], ms_theta: Sequence[torch.Tensor],
             max_phase: Sequence[float], ms_A: Optional[Sequence[torch.Tensor]] = None,
             dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """y = W (H_d + sum_i G_i diag(u_i) F_i), go.py:107-164; ms_theta[i] is
    [K_i] or [maps, K_i] (then y is [maps * C, N]).  dtype float64 (the
    oracle) or float32 (go.py's own precision, for the host baseline)."""
    cdt = torch.complex128 if dtype == torch.float64 else torch.complex64
    W = antenna_A.to(dtype) * torch.exp(-1j * antenna_theta.to(dtype))
    maps = ms_theta[0].reshape(-1, G_list[0].shape[1]).shape[0] if len(ms_theta) else 1
    out = []
    for m in range(maps):
        ch = H_d.to(cdt)
        for i in range(len(G_list)):
            th = ms_theta[i].reshape(maps, -1)[m].to(dtype)
            real_phase = torch.remainder(th, max_phase[i])  # go.py:131
            A = ms_A[i].to(dtype) if ms_A is not None else torch.ones_like(real_phase)
            u = A * torch.exp(1j * real_phase)
            ch = ch + torch.matmul(G_list[i].to(cdt) * u, F_list[i].to(cdt))
        out.append(torch.matmul(W, ch))
    return torch.cat(out, 0)


def received_chains(antenna_A, antenna_theta, channels: dict, ms_theta, max_phase, max_ms_reflex: int):
    """go.py:107-164 literally: every chain of metasurfaces of length
    1..max_ms_reflex without consecutive repeats (itertools.product), with
    ms{b}->ms{a} for b > a taken as the transpose of ms{a}->ms{b}."""
    import itertools
    W = antenna_A.to(torch.float64) * torch.exp(-1j * antenna_theta.to(torch.float64))
    ch = channels["src_to_dst"].to(torch.complex128)
    n = len(ms_theta)
    for L in range(1, max_ms_reflex + 1):
        for path in itertools.product(range(n), repeat=L):
            if any(path[i] == path[i + 1] for i in range(len(path) - 1)):
                continue
            m = channels["src_to_ms{%d}" % path[0]].to(torch.complex128)
            for i, ms_id in enumerate(path):
                u = torch.exp(1j * torch.remainder(ms_theta[ms_id].to(torch.float64), max_phase[ms_id]))
                m = m * u
                if i == len(path) - 1:
                    m = m @ channels["ms{%d}_to_dst" % ms_id].to(torch.complex128)
                else:
                    nxt = path[i + 1]
                    T = channels["ms{%d}_to_ms{%d}" % (ms_id, nxt)] if ms_id < nxt else \
                        channels["ms{%d}_to_ms{%d}" % (nxt, ms_id)].T
                    m = m @ T.to(torch.complex128)
            ch = ch + m
    return W @ ch


def loss_logsnr(Y: torch.Tensor, groups: int, n0: float) -> torch.Tensor:
    """-(1/N) sum_r log2(1 + max_c |y|^2 / n0) per group, summed."""
    GC, N = Y.shape
    s = (Y.real ** 2 + Y.imag ** 2).reshape(groups, GC // groups, N)
    best, _ = s.max(1)
    return (-torch.log2(1 + best / n0).mean(1)).sum()


def loss_capacity(Y: torch.Tensor, n0: float, w: float, rowmask: Optional[torch.Tensor]) -> torch.Tensor:
    """go.py:166-225 with loss_mode 'capacity' (one group)."""
    recv_amp = Y.abs()
    rx_rss = 20 * torch.log10(recv_amp + EPS32)
    rss_power = 10 ** (rx_rss / 10)
    cap = torch.log2(1 + rss_power / n0)
    rand_cap = cap.clone()
    if rowmask is not None:
        rand_cap[rowmask.bool()] = 0
    max_cap, _ = rand_cap.max(0)
    mask = max_cap != rand_cap
    masked_cap = rand_cap * mask
    return -w * max_cap.sum() + masked_cap.norm("fro")


def coverage_db(Y: torch.Tensor, groups: int = 1) -> torch.Tensor:
    GC, N = Y.shape
    p = (Y.real ** 2 + Y.imag ** 2).reshape(groups, GC // groups, N).amax(1)
    return (10 * torch.log10(p + 1e-30)).mean(1)
