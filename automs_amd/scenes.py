"""Seeded synthetic workloads of the channel and coverage path.

Workload names follow what they are:
  room_boxes      box room 10 x 8 x 3 m + 20 seeded boxes, 4x4 AP at 60 GHz,
                  64 x 64 Rx grid at z = 1.2 m (the plain, no-metasurface case)
  room_boxes_ms   the same room with one 32 x 32 metasurface on the x = 10 m
                  wall: traces AP->MS (G), MS->Rx (F) and AP->Rx with the
                  metasurface footprint absorbing (H_d)
  office          seeded office mesh (~100k triangles), 8x8 AP, 3 metasurfaces
                  of 48 x 48 elements, 256 x 256 Rx grid

All random draws come from splitmix64 so that the same seed gives the same
scene in every process.  Reflection magnitudes are drawn per surface (face)
and are constant over the 19 incidence bins of the reference's table
(go.py:683-691 reads one row of 19 bins per material).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .geometry import create_vertex_buffer, gen_ms, gen_pos_from_endpoint, gen_pos_matrix

MASK64 = (1 << 64) - 1
WAVE_LEN_60GHZ = 3e8 / 60e9


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) with 53-bit uniform doubles."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


@dataclass
class Scene:
    """Triangles [T, 9] f32 (v0, v1, v2), material per triangle, the 19-bin
    material table and an optional reflection phase per material (rad)."""
    tri: np.ndarray
    mat: np.ndarray
    mat_prop: np.ndarray
    refl_phase: Optional[np.ndarray] = None
    names: List[str] = field(default_factory=list)

    @property
    def n_tri(self) -> int:
        return int(self.tri.shape[0])

    def with_extra(self, tris: List, value: float, name: str, phase: float = 0.0) -> "Scene":
        """Append triangles with one new constant material (blackhole 0,
        reflector 1, door -0.7 in go.py:708-733)."""
        if not tris:
            return self
        t = np.asarray(tris, dtype=np.float32).reshape(-1, 9)
        mid = self.mat_prop.shape[0]
        rp = self.refl_phase if self.refl_phase is not None else np.zeros(mid, np.float32)
        return Scene(
            tri=np.concatenate([self.tri, t]),
            mat=np.concatenate([self.mat, np.full(len(t), mid, np.int32)]),
            mat_prop=np.concatenate([self.mat_prop, np.full((1, 19), value, np.float32)]),
            refl_phase=np.concatenate([rp, np.asarray([phase], np.float32)]),
            names=self.names + [name],
        )


@dataclass
class ArraySpec:
    pos: np.ndarray  # [n0, n1, 3] f32
    dir: np.ndarray  # [3]
    bptype: int
    radius: float

    @property
    def n(self) -> int:
        return int(self.pos.shape[0] * self.pos.shape[1])


@dataclass
class Workload:
    name: str
    scene: Scene
    tx: ArraySpec
    rx: ArraySpec
    wave_len: float
    max_reflection: int
    fib_n: int
    per_element: bool
    gridsize: int = 5
    ms: List[dict] = field(default_factory=list)
    extra: Dict = field(default_factory=dict)


def _faces(tris: List) -> np.ndarray:
    return np.asarray(tris, dtype=np.float32).reshape(-1, 9)


def room_boxes_scene(seed: int = 0, n_boxes: int = 20, room=(10.0, 8.0, 3.0), refl_phase: float = np.pi):
    """Box room (12 triangles) + n_boxes seeded boxes (12 each) resting on the
    floor; |Gamma| ~ U[0.3, 0.7] per face, reflection phase refl_phase."""
    rng = SplitMix64(seed)
    X, Y, Z = room
    tris = [_faces(create_vertex_buffer([X / 2, Y / 2, Z / 2], X, Y, Z))]
    for _ in range(n_boxes):
        sx, sy, sz = rng.uniform(0.3, 1.5), rng.uniform(0.3, 1.5), rng.uniform(0.3, 1.5)
        cx = rng.uniform(sx / 2, X - sx / 2)
        cy = rng.uniform(sy / 2, Y - sy / 2)
        tris.append(_faces(create_vertex_buffer([cx, cy, sz / 2], sx, sy, sz)))
    tri = np.concatenate(tris).astype(np.float32)
    n_faces = tri.shape[0] // 2
    gam = np.asarray([rng.uniform(0.3, 0.7) for _ in range(n_faces)], dtype=np.float32)
    mat = np.repeat(np.arange(n_faces, dtype=np.int32), 2)
    mat_prop = np.repeat(gam[:, None], 19, axis=1).astype(np.float32)
    rp = np.full(n_faces, refl_phase, dtype=np.float32)
    return Scene(tri=tri, mat=mat, mat_prop=mat_prop, refl_phase=rp, names=[f"face{i}" for i in range(n_faces)])


def room_boxes(seed: int = 0, fib_n: int = 1 << 20, n_rx=(64, 64), max_reflection: int = 3) -> Workload:
    """Plain workload: 4x4 half-wave AP at (0.05, 4.0, 2.5) facing +x (beam
    pattern 2), 64 x 64 isotropic Rx (radius 0.06 m) at z = 1.2 m, fib_n
    Fibonacci rays per AP element, <= 3 reflections."""
    wl = WAVE_LEN_60GHZ
    scene = room_boxes_scene(seed)
    tx = gen_pos_matrix([0.05, 4.0, 2.5], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0], (4, 4), wl / 2).numpy()
    rx = gen_pos_from_endpoint([0.1, 9.9, n_rx[0]], [0.1, 7.9, n_rx[1]], 1.2).numpy()
    return Workload(
        name="room_boxes", scene=scene,
        tx=ArraySpec(tx.astype(np.float32), np.array([1, 0, 0], np.float32), 2, 0.5e-5),
        rx=ArraySpec(rx.astype(np.float32), np.array([0, 0, 0], np.float32), 0, 0.06),
        wave_len=wl, max_reflection=max_reflection, fib_n=fib_n, per_element=True)


def room_boxes_ms(seed: int = 0, fib_n: int = 1 << 20, ms_fib_n: int = 1 << 24, n_rx=(64, 64),
                  ms_size=(32, 32), max_reflection: int = 3) -> Dict[str, Workload]:
    """Metasurface workload: returns the three traces {H_d, G, F} of one 32x32
    metasurface (half-wave pitch, cos^2 power pattern = beam pattern 1) on the
    x = 10 m wall, as build_channels does for one metasurface
    (go.py:1135-1224)."""
    base = room_boxes(seed, fib_n, n_rx, max_reflection)
    wl = base.wave_len
    pos, blackhole, normal = gen_ms([10.0, 4.0, 1.5], [-1.0, 0.0, 0.0], [0.0, 0.0, -1.0], ms_size, wl / 2)
    pos = pos.numpy().astype(np.float32)
    nrm = normal.numpy().astype(np.float32)
    ms = dict(pos=pos, blackhole=blackhole, normal=nrm, size=ms_size)
    hd_scene = base.scene.with_extra(blackhole, 0.0, "blackhole")
    out = {}
    out["H_d"] = Workload(name="room_boxes_ms/H_d", scene=hd_scene, tx=base.tx, rx=base.rx, wave_len=wl,
                          max_reflection=max_reflection, fib_n=fib_n, per_element=True, ms=[ms])
    out["G"] = Workload(name="room_boxes_ms/G", scene=base.scene, tx=base.tx,
                        rx=ArraySpec(pos, nrm, 1, 0.06), wave_len=wl, max_reflection=max_reflection,
                        fib_n=fib_n, per_element=True, ms=[ms])
    out["F"] = Workload(name="room_boxes_ms/F", scene=base.scene, tx=ArraySpec(pos, nrm, 1, 0.5e-5),
                        rx=base.rx, wave_len=wl, max_reflection=max_reflection, fib_n=ms_fib_n,
                        per_element=False, ms=[ms])
    return out
