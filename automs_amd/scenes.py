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
    info: Dict = field(default_factory=dict)  # generator facts (e.g. the office's door gaps)

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
            info=self.info,
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


def _tess_quad(origin, u, v, nu: int, nv: int) -> np.ndarray:
    """Planar rectangle origin + s u + t v (s, t in [0, 1]) as nu x nv cells of
    two triangles each: [2 nu nv, 9] f32 (all exactly coplanar)."""
    o, u, v = (np.asarray(x, np.float64) for x in (origin, u, v))
    s = np.arange(nu + 1) / nu
    t = np.arange(nv + 1) / nv
    P = o[None, None, :] + s[:, None, None] * u[None, None, :] + t[None, :, None] * v[None, None, :]
    a, b = P[:-1, :-1], P[1:, :-1]
    c, d = P[1:, 1:], P[:-1, 1:]
    t1 = np.concatenate([a, b, c], -1).reshape(-1, 9)
    t2 = np.concatenate([a, c, d], -1).reshape(-1, 9)
    return np.concatenate([t1, t2]).astype(np.float32)


def _box(cx, cy, cz, sx, sy, sz) -> np.ndarray:
    return _faces(create_vertex_buffer([cx, cy, cz], sx, sy, sz))


def office_scene(seed: int = 0, size=(24.0, 16.0, 3.0), cell: float = 0.15, refl_phase: float = np.pi):
    """Seeded office: tessellated floor / ceiling / outer walls, a corridor
    along y with 6 rooms on each side (interior walls 0.12 m thick with door
    gaps), and per room a grid of desk clusters (desk top, two side panels, a
    monitor, a chair of seat + back + four legs) and a bookshelf of
    shelves.  Materials: floor, ceiling, outer wall, inner wall, desk, chair,
    monitor, shelf, each |Gamma| ~ U[0.3, 0.7], constant over the 19 bins.
    About 1e5 triangles at cell = 0.15 m."""
    rng = SplitMix64(seed)
    X, Y, Z = size
    groups = {}  # material name -> list of triangle arrays
    gaps = []  # (corridor wall y, gap x0, gap x1)

    def add(name, t):
        groups.setdefault(name, []).append(t)

    nx, ny, nz = int(round(X / cell)), int(round(Y / cell)), int(round(Z / cell))
    add("floor", _tess_quad([0, 0, 0], [X, 0, 0], [0, Y, 0], nx, ny))
    add("ceiling", _tess_quad([0, 0, Z], [X, 0, 0], [0, Y, 0], nx, ny))
    add("outer_wall", _tess_quad([0, 0, 0], [X, 0, 0], [0, 0, Z], nx, nz))
    add("outer_wall", _tess_quad([0, Y, 0], [X, 0, 0], [0, 0, Z], nx, nz))
    add("outer_wall", _tess_quad([0, 0, 0], [0, Y, 0], [0, 0, Z], ny, nz))
    add("outer_wall", _tess_quad([X, 0, 0], [0, Y, 0], [0, 0, Z], ny, nz))
    cy0, cy1 = Y / 2 - 1.0, Y / 2 + 1.0  # corridor
    th = 0.12
    rooms = 6
    rw = X / rooms
    for side, (y0, y1) in enumerate(((0.0, cy0), (cy1, Y))):
        wall_y = cy0 if side == 0 else cy1
        for r in range(rooms):
            x0 = r * rw
            # corridor wall of this room with a 1 m door gap at a seeded spot
            gx = x0 + rng.uniform(0.5, rw - 1.5)
            gaps.append((wall_y, gx, gx + 1.0))
            for a, b in ((x0, gx), (gx + 1.0, x0 + rw)):
                if b - a > 0.05:
                    add("inner_wall", _box((a + b) / 2, wall_y, Z / 2, b - a, th, Z))
            if r > 0:  # wall between rooms
                add("inner_wall", _box(x0, (y0 + y1) / 2, Z / 2, th, y1 - y0, Z))
            # desk clusters on a 2 x 3 grid
            for i in range(2):
                for j in range(3):
                    dx = x0 + 0.8 + i * (rw - 1.6) + rng.uniform(-0.2, 0.2)
                    dy = y0 + (j + 0.5) * (y1 - y0) / 3 + rng.uniform(-0.2, 0.2)
                    w_, d_ = rng.uniform(1.2, 1.6), rng.uniform(0.6, 0.8)
                    add("desk", _box(dx, dy, 0.74, w_ * 0.5, d_, 0.03))
                    add("desk", _box(dx - w_ * 0.23, dy, 0.36, 0.03, d_, 0.72))
                    add("desk", _box(dx + w_ * 0.23, dy, 0.36, 0.03, d_, 0.72))
                    add("monitor", _box(dx, dy + d_ * 0.3, 0.98, 0.55, 0.04, 0.35))
                    cxh, cyh = dx, dy - d_ * 0.5 - 0.3
                    add("chair", _box(cxh, cyh, 0.46, 0.45, 0.45, 0.05))
                    add("chair", _box(cxh, cyh - 0.21, 0.75, 0.45, 0.04, 0.55))
                    for lx in (-0.19, 0.19):
                        for ly in (-0.19, 0.19):
                            add("chair", _box(cxh + lx, cyh + ly, 0.22, 0.03, 0.03, 0.44))
            # bookshelf against the outer wall: frame + 5 shelves
            bx = x0 + rng.uniform(0.8, rw - 0.8)
            by = (0.25 if side == 0 else Y - 0.25)
            add("shelf", _box(bx - 0.45, by, 1.0, 0.03, 0.35, 2.0))
            add("shelf", _box(bx + 0.45, by, 1.0, 0.03, 0.35, 2.0))
            for k in range(6):
                add("shelf", _box(bx, by, 0.05 + 0.38 * k, 0.9, 0.35, 0.03))
    names = list(groups.keys())
    tri = np.concatenate([np.concatenate(groups[n]) for n in names]).astype(np.float32)
    mat = np.concatenate([np.full(sum(len(t) for t in groups[n]), i, np.int32) for i, n in enumerate(names)])
    gam = np.asarray([rng.uniform(0.3, 0.7) for _ in names], np.float32)
    mat_prop = np.repeat(gam[:, None], 19, axis=1).astype(np.float32)
    return Scene(tri=tri, mat=mat, mat_prop=mat_prop, refl_phase=np.full(len(names), refl_phase, np.float32),
                 names=names, info=dict(door_gaps=gaps, corridor=(cy0, cy1), wall_thickness=th))


def _on_wall(x: float, wall_y: float, gaps, clear: float = 0.4) -> float:
    """x moved along a corridor wall until it is at least `clear` from every
    door gap of that wall (the nearest such position)."""
    spans = sorted((a, b) for (y, a, b) in gaps if abs(y - wall_y) < 1e-6)
    for a, b in spans:
        if a - clear < x < b + clear:
            x = a - clear if x - (a - clear) < (b + clear) - x else b + clear
    return x


def office(seed: int = 0, fib_n: int = 1 << 24, ms_fib_n: int = 1 << 24, n_rx=(256, 256), ap=(8, 8),
           ms_size=(48, 48), max_reflection: int = 5, cell: float = 0.15) -> Dict[str, Workload]:
    """Office workload: 8x8 AP (beam pattern 2) on the corridor's west end,
    three 48 x 48 metasurfaces that the AP sees along the corridor (on the
    south and north corridor walls, clear of their door gaps, facing the
    corridor, and on the east wall at the corridor's end), 256 x 256 Rx at
    z = 1.2 m.  Traces: H_d, G{i} (AP -> metasurface i), F{i} (metasurface i
    -> Rx) and T{i}{j} (metasurface i -> j, i < j), as build_channels forms
    them; OFFICE_CHANNEL_KEYS gives their go.py channel names.
    fib_n rays in total per AP trace (distributed over the elements) and
    ms_fib_n per metasurface trace.  Returns H_d and G{i}, F{i} traces."""
    wl = WAVE_LEN_60GHZ
    scene = office_scene(seed, cell=cell)
    X, Y, Z = 24.0, 16.0, 3.0
    tx = gen_pos_matrix([0.3, Y / 2, 2.4], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0], ap, wl / 2).numpy()
    rx = gen_pos_from_endpoint([0.1, X - 0.1, n_rx[0]], [0.1, Y - 0.1, n_rx[1]], 1.2).numpy()
    tx_spec = ArraySpec(tx.astype(np.float32), np.array([1, 0, 0], np.float32), 2, 0.5e-5)
    rx_spec = ArraySpec(rx.astype(np.float32), np.array([0, 0, 0], np.float32), 0, 0.06)
    # metasurfaces: on the corridor faces of its two walls and on the east wall
    cy0, cy1 = scene.info["corridor"]
    half = scene.info["wall_thickness"] / 2
    gaps = scene.info["door_gaps"]
    placements = [([_on_wall(10.0, cy0, gaps), cy0 + half, 2.0], [0.0, 1.0, 0.0]),
                  ([_on_wall(16.0, cy1, gaps), cy1 - half, 2.0], [0.0, -1.0, 0.0]),
                  ([X, Y / 2, 1.8], [-1.0, 0.0, 0.0])]
    ms_list = []
    for c, k in placements:
        pos, bh, nrm = gen_ms(c, k, [0.0, 0.0, -1.0], ms_size, wl / 2)
        ms_list.append(dict(pos=pos.numpy().astype(np.float32), blackhole=bh, normal=nrm.numpy().astype(np.float32),
                            size=ms_size))
    out = {}
    all_bh = [t for m in ms_list for t in m["blackhole"]]
    out["H_d"] = Workload("office/H_d", scene.with_extra(all_bh, 0.0, "blackhole"), tx_spec, rx_spec, wl,
                          max_reflection, fib_n, False, ms=ms_list)
    for i, m in enumerate(ms_list):
        others = [t for j, mm in enumerate(ms_list) if j != i for t in mm["blackhole"]]
        sc = scene.with_extra(others, 0.0, "blackhole")
        out[f"G{i}"] = Workload(f"office/G{i}", sc, tx_spec, ArraySpec(m["pos"], m["normal"], 1, 0.06), wl,
                                max_reflection, fib_n, False, ms=ms_list)
        out[f"F{i}"] = Workload(f"office/F{i}", sc, ArraySpec(m["pos"], m["normal"], 1, 0.5e-5), rx_spec, wl,
                                max_reflection, ms_fib_n, False, ms=ms_list)
    # metasurface i -> j (i < j), the others absorbing (go.py:1226-1245)
    for i in range(len(ms_list)):
        for j in range(i + 1, len(ms_list)):
            others = [t for k, mm in enumerate(ms_list) if k not in (i, j) for t in mm["blackhole"]]
            sc = scene.with_extra(others, 0.0, "blackhole")
            out[f"T{i}{j}"] = Workload(f"office/T{i}{j}", sc, ArraySpec(ms_list[i]["pos"], ms_list[i]["normal"], 1,
                                                                       0.5e-5),
                                       ArraySpec(ms_list[j]["pos"], ms_list[j]["normal"], 1, 0.06), wl,
                                       max_reflection, ms_fib_n, False, ms=ms_list)
    return out


OFFICE_CHANNEL_KEYS = {"H_d": "src_to_dst", "G0": "src_to_ms{0}", "G1": "src_to_ms{1}", "G2": "src_to_ms{2}",
                       "F0": "ms{0}_to_dst", "F1": "ms{1}_to_dst", "F2": "ms{2}_to_dst",
                       "T01": "ms{0}_to_ms{1}", "T02": "ms{0}_to_ms{2}", "T12": "ms{1}_to_ms{2}"}
