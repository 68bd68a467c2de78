"""Drop-in replacements for the reference's Python-facing channel API.

* ``torch.ops.pathtrace.torch_launch_pathtrace`` is registered with the
  argument list and order of optixPathTracer.cpp:781-791 and accumulates into
  ``result`` like the reference (go.py:671-677, :742-765).
* ``build_channel`` / ``build_channels`` follow go.py:623-767 and
  go.py:1135-1247: scene folders of binary STL files whose material is the
  file-name suffix, mat.txt (19 reflectivity bins per material), synthetic
  blackhole / reflector / door materials (0, 1, -0.7).

The op traces the reference's own ray set: launch_x x launch_y threads x
samples_per_launch TEA/LCG rays from random Tx elements with the per-thread
min-hit gating and reserve_size records (optixPathTracer.cu:142-190).
Differences from the reference, all deliberate:
* the AABB buffers are accepted and ignored: spheres are defined by the
  element positions and src_r / dst_r;
* the scene hierarchy is built once per distinct (triangles, materials,
  table) and kept for later calls (a small cache of scene handles), where
  the reference rebuilds its acceleration structures on every call
  (optixPathTracer.cpp:834-840);
* failures raise instead of being printed and swallowed
  (optixPathTracer.cpp:858-860).
"""
from __future__ import annotations

import math
import os
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from .scenes import ArraySpec, Scene
from .tracer import TraceScene, trace, trace_many

# ---------------------------------------------------------------- scene I/O --

_STL_DTYPE = np.dtype([("normal", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")])


def read_stl(path: str) -> np.ndarray:
    """Binary STL -> triangles [T, 3, 3] f32 (what numpy-stl's Mesh.vectors
    gives go.py:700)."""
    with open(path, "rb") as f:
        head = f.read(84)
        if len(head) < 84:
            raise ValueError(f"{path}: not a binary STL")
        n = int(np.frombuffer(head[80:84], "<u4")[0])
        data = np.frombuffer(f.read(50 * n), dtype=_STL_DTYPE, count=n)
    return np.ascontiguousarray(data["v"], dtype=np.float32)


def read_mat(scene_folder: str) -> Tuple[List[str], np.ndarray]:
    """mat.txt: 'name v0 .. v18' per line (go.py:681-691)."""
    names, rows = [], []
    with open(os.path.join(scene_folder, "mat.txt")) as f:
        for line in f:
            vals = line.split()
            if not vals:
                continue
            names.append(vals[0])
            rows.append([float(x) for x in vals[1:]])
    return names, np.asarray(rows, dtype=np.float32)


def get_walls(scene_folder: str) -> List[List[List[float]]]:
    """wall.txt: per wall a name line and 3 vertex lines (go.py:597-620)."""
    walls = []
    with open(os.path.join(scene_folder, "wall.txt")) as f:
        lines = f.readlines()
    for i in range(len(lines) // 4):
        walls.append([[float(x) for x in lines[4 * i + 1 + j].split()] for j in range(3)])
    return walls


def get_ap(scene_folder: str):
    """ap.txt: centre line and direction line (go.py:533-551)."""
    with open(os.path.join(scene_folder, "ap.txt")) as f:
        lines = f.readlines()
    return [float(x) for x in lines[0].split()], [float(x) for x in lines[1].split()]


def get_doors(scene_folder: str, angle: float):
    """door.txt: 3 vertex lines + an opening-side flag per door; the door leaf
    is rotated by angle (rad) about its hinge (go.py:554-594)."""
    import copy
    doors = []
    with open(os.path.join(scene_folder, "door.txt")) as f:
        lines = f.readlines()
    for i in range(len(lines) // 4):
        tri = [[float(x) for x in lines[4 * i + j].split()] for j in range(3)]
        theta0 = math.atan2(tri[2][1] - tri[1][1], tri[2][0] - tri[1][0])
        theta = theta0 + angle * (1 - 2 * int(lines[4 * i + 3]))
        tri[2][0] = tri[1][0] + math.cos(theta)
        tri[2][1] = tri[1][1] + math.sin(theta)
        t2 = copy.deepcopy(tri)
        t2[1][2] = t2[0][2]
        t2[1][0] = t2[2][0]
        t2[1][1] = t2[2][1]
        doors.append([tri, t2])
    return doors


def load_scene_folder(scene_folder: str, blackhole_buffer=(), reflector_buffer=(), door_buffer=(),
                      refl_phase: float = 0.0) -> Scene:
    """Scene assembly of go.py:679-737: STL files (material = file-name
    suffix) plus synthetic blackhole (0), reflector (1) and door (-0.7)."""
    names, mat_prop = read_mat(scene_folder)
    tris, mats = [], []
    for fn in sorted(os.listdir(scene_folder)):
        if not fn.endswith(".stl"):
            continue
        mat_name = fn.split("_")[-1].split(".stl")[0]
        mid = names.index(mat_name)
        t = read_stl(os.path.join(scene_folder, fn))
        tris.append(t.reshape(-1, 9))
        mats.append(np.full(len(t), mid, np.int32))
    tri = np.concatenate(tris) if tris else np.zeros((0, 9), np.float32)
    mat = np.concatenate(mats) if mats else np.zeros(0, np.int32)
    sc = Scene(tri=tri, mat=mat, mat_prop=mat_prop, refl_phase=np.full(len(names), refl_phase, np.float32),
               names=list(names))
    sc = sc.with_extra([t for b in blackhole_buffer for t in b], 0.0, "blackhole", refl_phase)
    sc = sc.with_extra([t for b in reflector_buffer for t in b], 1.0, "reflector", refl_phase)
    sc = sc.with_extra([t for b in door_buffer for t in b], -0.7, "door", refl_phase)
    return sc


# ------------------------------------------------------------- torch op -----

def _np(t) -> np.ndarray:
    return t.detach().to("cpu", torch.float32).contiguous().numpy() if isinstance(t, torch.Tensor) \
        else np.asarray(t, np.float32)


def torch_launch_pathtrace(result: torch.Tensor, src: torch.Tensor, dst: torch.Tensor,
                           src_aabb_buffer: torch.Tensor, dst_aabb_buffer: torch.Tensor,
                           scene_triangle_buffer: torch.Tensor, mat: torch.Tensor, samples_per_launch: int,
                           launch_x: int, launch_y: int, wave_len: float, max_reflection: int, reserve_size: int,
                           mat_prop: torch.Tensor, src_dir: torch.Tensor, dst_dir: torch.Tensor,
                           optix_log_level: int, src_bptype: int, dst_bptype: int, dst_girdsize: int,
                           src_r: float, dst_r: float) -> None:
    """optixPathTracer.cpp:781-861 on the MI355X tracer."""
    tri = _np(scene_triangle_buffer).reshape(-1, 9)
    scene = Scene(tri=tri, mat=mat.detach().to("cpu", torch.int32).contiguous().numpy(),
                  mat_prop=_np(mat_prop).reshape(-1, 19))
    s, d = _np(src), _np(dst)
    if s.ndim != 3 or d.ndim != 3:
        raise ValueError("src and dst must be [n0, n1, 3] (optixPathTracer.cpp:814-821)")
    tx = ArraySpec(s, _np(src_dir).reshape(3), int(src_bptype), float(src_r))
    rx = ArraySpec(d, _np(dst_dir).reshape(3), int(dst_bptype), float(dst_r))
    ts = _cached_scene(scene)
    _, st, _ = trace(ts, tx, rx, wave_len, int(max_reflection), 0, per_element=False,
                     gridsize=int(dst_girdsize), result=result,
                     reference_launch=(launch_x, launch_y, samples_per_launch, reserve_size))
    if optix_log_level > 0:
        print(f"rays {st['rays']}, unique records {st['records_unique']}, paths {st['paths']}, "
              f"tasks {st['tasks']} in {st['rounds']} rounds, chains {st['chains']} "
              f"({100.0 * st['successes'] / max(1, st['chains']):.2f}% valid), {st['t_total_ms']:.1f} ms")


_SCENE_CACHE: "OrderedDict[bytes, TraceScene]" = None
_SCENE_CACHE_MAX = 4


def _cached_scene(scene: Scene) -> TraceScene:
    """The scene handle of these triangles, materials and table: built on
    first use, kept for the next calls (at most _SCENE_CACHE_MAX scenes, the
    least recently used closed first).  One handle serves one trace at a
    time, as the op's callers (one Python thread) trace."""
    import hashlib
    from collections import OrderedDict
    global _SCENE_CACHE
    if _SCENE_CACHE is None:
        _SCENE_CACHE = OrderedDict()
    h = hashlib.blake2b(digest_size=20)
    for a in (scene.tri, scene.mat, scene.mat_prop):
        a = np.ascontiguousarray(a)
        h.update(str((a.dtype, a.shape)).encode())
        h.update(a.tobytes())
    key = h.digest()
    ts = _SCENE_CACHE.get(key)
    if ts is None:
        ts = TraceScene(scene)
        _SCENE_CACHE[key] = ts
        while len(_SCENE_CACHE) > _SCENE_CACHE_MAX:
            _, old = _SCENE_CACHE.popitem(last=False)
            old.close()
    else:
        _SCENE_CACHE.move_to_end(key)
    return ts


_OP_SCHEMA = ("torch_launch_pathtrace(Tensor(a!) result, Tensor src, Tensor dst, Tensor src_aabb_buffer, "
              "Tensor dst_aabb_buffer, Tensor scene_triangle_buffer, Tensor mat, int samples_per_launch, "
              "int launch_x, int launch_y, float wave_len, int max_reflection, int reserve_size, "
              "Tensor mat_prop, Tensor src_dir, Tensor dst_dir, int optix_log_level, int src_bptype, "
              "int dst_bptype, int dst_girdsize, float src_r, float dst_r) -> ()")
_registered = False


def register_torch_op() -> None:
    """Register torch.ops.pathtrace.torch_launch_pathtrace (the reference
    registers it in TORCH_LIBRARY(pathtrace), optixPathTracer.cpp:868-870)."""
    global _registered, _torch_lib
    if _registered:
        return
    _torch_lib = torch.library.Library("pathtrace", "DEF")
    _torch_lib.define(_OP_SCHEMA)
    _torch_lib.impl("torch_launch_pathtrace", torch_launch_pathtrace, "CompositeExplicitAutograd")
    _registered = True


def init(op_path: Optional[str] = None, seed: int = 5) -> None:
    """go.py:514-530: make the op available and seed torch."""
    register_torch_op()
    torch.manual_seed(seed)


# -------------------------------------------------------- build_channel -----

def _channel_job(src, dst, src_dir, dst_dir, src_bptype: int, dst_bptype: int, wave_len: float,
                 scene_folder: Optional[str] = None, max_reflection: int = 4, dst_radius: float = 0.06,
                 src_radius: float = 0.5e-5, samples_per_launch: int = 1, reserved_size: int = 100,
                 blackhole_buffer=(), reflector_buffer=(), door_buffer=(), dst_gridsize: int = 5,
                 scene: Optional[Scene] = None, rays: Optional[int] = None, per_element: bool = False,
                 refl_phase: float = 0.0) -> Tuple[Scene, dict]:
    """The scene and the trace() arguments (without the handle) of one
    build_channel call."""
    if scene is None:
        if scene_folder is None:
            raise ValueError("build_channel needs scene_folder or scene")
        scene = load_scene_folder(scene_folder, blackhole_buffer, reflector_buffer, door_buffer, refl_phase)
    else:
        scene = scene.with_extra([t for b in blackhole_buffer for t in b], 0.0, "blackhole", refl_phase)
        scene = scene.with_extra([t for b in reflector_buffer for t in b], 1.0, "reflector", refl_phase)
        scene = scene.with_extra([t for b in door_buffer for t in b], -0.7, "door", refl_phase)
    s, d = _np(src), _np(dst)
    tx = ArraySpec(s, _np(src_dir).reshape(3), int(src_bptype), float(src_radius))
    rx = ArraySpec(d, _np(dst_dir).reshape(3), int(dst_bptype), float(dst_radius))
    if rays is None:
        job = dict(tx=tx, rx=rx, wave_len=wave_len, max_reflection=int(max_reflection), fib_n=0, per_element=False,
                   gridsize=int(dst_gridsize),
                   reference_launch=(100, 100, int(samples_per_launch), int(reserved_size)))
    else:
        job = dict(tx=tx, rx=rx, wave_len=wave_len, max_reflection=int(max_reflection), fib_n=int(rays),
                   per_element=per_element, gridsize=int(dst_gridsize))
    return scene, job


def build_channel(src: torch.Tensor, dst: torch.Tensor, src_dir, dst_dir, src_bptype: int, dst_bptype: int,
                  wave_len: float, scene_folder: Optional[str] = None, max_reflection: int = 4,
                  dst_radius: float = 0.06, src_radius: float = 0.5e-5, samples_per_launch: int = 1,
                  reserved_size: int = 100, blackhole_buffer=(), reflector_buffer=(), door_buffer=(),
                  dst_gridsize: int = 5, scene: Optional[Scene] = None, rays: Optional[int] = None,
                  per_element: bool = False, refl_phase: float = 0.0, device: str = "cuda") -> torch.Tensor:
    """go.py:623-767: trace src -> dst and return complex64 [n_src, n_dst].

    Either scene_folder (STL files + mat.txt) or an assembled `scene`.
    rays None: the reference's own launch (go.py:739-765), 100 x 100 threads
    x samples_per_launch TEA/LCG rays with reserved_size records per thread;
    rays given: that many Fibonacci rays (per src element if per_element)."""
    sc, job = _channel_job(src, dst, src_dir, dst_dir, src_bptype, dst_bptype, wave_len, scene_folder,
                           max_reflection, dst_radius, src_radius, samples_per_launch, reserved_size,
                           blackhole_buffer, reflector_buffer, door_buffer, dst_gridsize, scene, rays, per_element,
                           refl_phase)
    result = torch.zeros(job["tx"].n, job["rx"].n, 2, dtype=torch.float32, device=device)
    ts = TraceScene(sc)
    try:
        trace(ts, result=result, device=device, **job)
    finally:
        ts.close()
    return torch.complex(result[..., 0], result[..., 1])


def build_channels(src: torch.Tensor, dst: torch.Tensor, ms_list, ap_dir, scene_folder: Optional[str],
                   wave_len: float, max_reflection: int, reflector_buffer=(), door_buffer=(),
                   scene: Optional[Scene] = None, rays: Optional[int] = None, ms_rays: Optional[int] = None,
                   per_element: bool = False, refl_phase: float = 0.0) -> Dict[str, torch.Tensor]:
    """go.py:1135-1247: src->dst with every metasurface absorbing, then
    src->ms{i}, ms{i}->dst and ms{i}->ms{j} with the other metasurfaces
    absorbing.  ms_list items need pos, send_dir, recv_dir, blackhole."""
    dir_no = torch.zeros(3)
    ap_dir = torch.as_tensor(ap_dir, dtype=torch.float32)
    n = len(ms_list)
    kw = dict(scene_folder=scene_folder, max_reflection=max_reflection, reflector_buffer=reflector_buffer,
              door_buffer=door_buffer, scene=scene, refl_phase=refl_phase)
    spec = {"src_to_dst": _channel_job(src, dst, ap_dir, dir_no, 2, 0, wave_len,
                                       blackhole_buffer=[m.blackhole for m in ms_list], rays=rays,
                                       per_element=per_element, **kw)}
    for i in range(n):
        others = [ms_list[k].blackhole for k in range(n) if k != i]
        spec["src_to_ms{%d}" % i] = _channel_job(src, ms_list[i].pos, ap_dir, ms_list[i].recv_dir, 2, 1, wave_len,
                                                 blackhole_buffer=others, rays=rays, per_element=per_element,
                                                 **kw)
        spec["ms{%d}_to_dst" % i] = _channel_job(ms_list[i].pos, dst, ms_list[i].send_dir, dir_no, 1, 0, wave_len,
                                                 blackhole_buffer=others, rays=ms_rays or rays, **kw)
        for j in range(i + 1, n):
            oth = [ms_list[k].blackhole for k in range(n) if k not in (i, j)]
            spec["ms{%d}_to_ms{%d}" % (i, j)] = _channel_job(ms_list[i].pos, ms_list[j].pos, ms_list[i].send_dir,
                                                             ms_list[j].recv_dir, 1, 1, wave_len,
                                                             blackhole_buffer=oth, rays=ms_rays or rays, **kw)
    # every channel set is an independent trace with its own scene: build the
    # scene hierarchies on host threads and trace the sets concurrently
    from concurrent.futures import ThreadPoolExecutor

    names = list(spec)
    with ThreadPoolExecutor(max_workers=min(8, len(names))) as ex:
        futures = [ex.submit(TraceScene, spec[k][0]) for k in names]
    handles = []
    err = None
    for f in futures:  # keep every handle that was built, so that all of them are closed
        try:
            handles.append(f.result())
        except BaseException as e:
            err = err or e
    if err is not None:
        for h in handles:
            h.close()
        raise err
    try:
        res = trace_many([dict(ts=h, **spec[k][1]) for k, h in zip(names, handles)], max_concurrent=4)
    finally:
        torch.cuda.synchronize()
        for h in handles:
            h.close()
    return {k: torch.complex(r[0][..., 0], r[0][..., 1]) for k, r in zip(names, res)}
