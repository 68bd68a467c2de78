"""Element-grid and placement helpers with the signatures of the reference's
geometry generators (go.py:248-293, :339-511, :1309-1338).

They build the same positions as the reference; the device defaults to the
CPU instead of the reference's hard-coded "cuda" because the tracer takes host
positions.
"""
from __future__ import annotations

import math
from typing import Literal, Tuple

import torch
import torch.nn.functional as F


def _t(x, device="cpu"):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float32)
    return torch.tensor(x, dtype=torch.float32, device=device)


def gen_direction_from_angle(theta: float, phi: float, up, polar_axis,
                             unit: Literal["deg", "rad"] = "deg", device="cpu") -> torch.Tensor:
    """Direction from polar/azimuth angles in the (polar_axis, up x polar_axis)
    frame, go.py:248-293."""
    if unit == "deg":
        theta = theta / 180 * math.pi
        phi = phi / 180 * math.pi
    up = _t(up, device)
    polar_axis = _t(polar_axis, device)
    i_direction = F.normalize(polar_axis, dim=-1)
    j_direction = F.normalize(torch.linalg.cross(up, i_direction), dim=-1)
    k_direction = torch.linalg.cross(i_direction, j_direction)
    ijk = torch.tensor([math.cos(theta) * math.cos(phi), math.cos(theta) * math.sin(phi), math.sin(theta)],
                       dtype=torch.float32, device=device)
    return ijk @ torch.stack([i_direction, j_direction, k_direction], dim=0)


def gen_pos_matrix(center, k_direction, i_direction, size: Tuple[int, int], spacing: float,
                   device="cpu") -> torch.Tensor:
    """Array element positions [size_i, size_j, 3], go.py:339-385."""
    size_i, size_j = size
    center = _t(center, device)
    i_direction = _t(i_direction, device)
    k_direction = _t(k_direction, device)
    j_direction = torch.linalg.cross(k_direction, i_direction)
    i_direction = F.normalize(i_direction, dim=-1)
    j_direction = F.normalize(j_direction, dim=-1)
    k_direction = F.normalize(k_direction, dim=-1)
    pos_i = torch.linspace(-(size_i - 1) / 2, (size_i - 1) / 2, size_i, device=device)
    pos_j = torch.linspace(-(size_j - 1) / 2, (size_j - 1) / 2, size_j, device=device)
    pos_i, pos_j = torch.meshgrid([pos_i, pos_j], indexing="ij")
    pos_ijk = torch.stack([pos_i.flatten(), pos_j.flatten(), torch.zeros(size_i * size_j, device=device)], -1)
    transform = torch.stack([i_direction, j_direction, k_direction], dim=0)
    pos = (pos_ijk * spacing) @ transform + center
    return pos.reshape(tuple(size) + (3,))


def gen_ms(center, k_direction, i_direction, size: Tuple[int, int], spacing: float, thickness: float = 0.001,
           device="cpu"):
    """Metasurface element grid, its blackhole footprint (2 triangles) and its
    normal, go.py:388-443."""
    size_i, size_j = size
    center = _t(center, device).clone()
    i_direction = _t(i_direction, device)
    k_direction = _t(k_direction, device)
    j_direction = torch.linalg.cross(k_direction, i_direction)
    i_direction = F.normalize(i_direction, dim=-1)
    j_direction = F.normalize(j_direction, dim=-1)
    k_direction = F.normalize(k_direction, dim=-1)
    center = center + k_direction * thickness
    pos = gen_pos_matrix(center, k_direction, i_direction, size, spacing, device=device)
    i_off = i_direction * size_i / 2 * spacing
    j_off = j_direction * size_j / 2 * spacing
    ul = (center - i_off - j_off).tolist()
    dr = (center + i_off + j_off).tolist()
    ur = (center + i_off - j_off).tolist()
    dl = (center - i_off + j_off).tolist()
    blackhole = [[ul, ur, dl], [dl, dr, ur]]
    return pos, blackhole, k_direction


def gen_sth_from_wall(wall, ij: Tuple[float, float], device="cpu"):
    """Centre, normal and down direction of a point on a wall given as
    [v_upleft, v_downleft, v_downright], go.py:446-477."""
    up, left = ij
    v_ul, v_dl, v_dr = (_t(v, device) for v in wall)
    v_ur = v_ul + v_dr - v_dl
    v_up = (1 - left) * v_ul + left * v_ur
    v_down = (1 - left) * v_dl + left * v_dr
    v = up * v_down + (1 - up) * v_up
    i_direction = F.normalize(v_dl - v_ul, dim=-1)
    j_direction = F.normalize(v_dr - v_dl, dim=-1)
    k_direction = torch.linalg.cross(i_direction, j_direction)
    return v, k_direction, i_direction


def gen_pos_from_endpoint(x, y, z, device="cpu") -> torch.Tensor:
    """Regular grid from [start, end, num] per axis, go.py:480-511."""
    pos_list, num_list = [], []
    for arg in (x, y, z):
        if not isinstance(arg, (list, tuple)):
            arg = [arg, arg, 1]
        start, end, num = arg
        num_list.append(num)
        pos_list.append(torch.linspace(start, end, num, dtype=torch.float32, device=device))
    num_list.remove(1)
    num_list += [-1]
    res = torch.meshgrid(pos_list, indexing="ij")
    return torch.stack(res, dim=-1).reshape(num_list)


_BOX_TEMPLATE = [
    [[-1, -1, 1], [1, -1, 1], [1, 1, 1]],  # top
    [[-1, -1, 1], [-1, 1, 1], [1, 1, 1]],
    [[-1, -1, -1], [1, -1, -1], [1, 1, -1]],  # bottom
    [[-1, -1, -1], [-1, 1, -1], [1, 1, -1]],
    [[-1, -1, -1], [-1, -1, 1], [-1, 1, 1]],  # left
    [[-1, -1, -1], [-1, 1, -1], [-1, 1, 1]],
    [[1, -1, -1], [1, -1, 1], [1, 1, 1]],  # right
    [[1, -1, -1], [1, 1, -1], [1, 1, 1]],
    [[-1, 1, -1], [1, 1, -1], [1, 1, 1]],  # front
    [[-1, 1, -1], [-1, 1, 1], [1, 1, 1]],
    [[-1, -1, -1], [1, -1, -1], [1, -1, 1]],  # back
    [[-1, -1, -1], [-1, -1, 1], [1, -1, 1]],
]


def create_vertex_buffer(center, xlen: float = 0.2, ylen: float = 0.2, zlen: float = 1.5):
    """12 triangles of an axis-aligned box (2 per face), go.py:1309-1338."""
    import numpy as np

    t = np.asarray(_BOX_TEMPLATE, dtype=np.float32) / 2 * np.asarray([xlen, ylen, zlen], dtype=np.float32)
    return (t + np.asarray(center, dtype=np.float32)).tolist()
