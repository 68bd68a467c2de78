"""Host API of the channel tracer (the counterpart of go.py's build_channel /
build_channels and of the torch op torch.ops.pathtrace.torch_launch_pathtrace,
optixPathTracer.cpp:781-870).  Compute runs in libams.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np
import torch

from . import _lib
from .scenes import ArraySpec, Scene, Workload


class TraceScene:
    """A scene handle: triangles, materials and the BVH built once on the
    host and kept in device memory across traces."""

    def __init__(self, scene: Scene):
        L = _lib.lib()
        self.scene = scene
        tri = np.ascontiguousarray(scene.tri, dtype=np.float32).reshape(-1)
        mat = np.ascontiguousarray(scene.mat, dtype=np.int32)
        mp = np.ascontiguousarray(scene.mat_prop, dtype=np.float32).reshape(-1)
        rp = None if scene.refl_phase is None else np.ascontiguousarray(scene.refl_phase, dtype=np.float32)
        h = C.c_void_p()
        fp = C.POINTER(C.c_float)
        _lib.check(L.ams_scene_create(
            tri.ctypes.data_as(fp), int(scene.n_tri), mat.ctypes.data_as(C.POINTER(C.c_int32)),
            int(scene.mat_prop.shape[0]), mp.ctypes.data_as(fp),
            rp.ctypes.data_as(fp) if rp is not None else None, C.byref(h)), "ams_scene_create")
        self.handle = h

    def plane_ids(self) -> np.ndarray:
        out = np.zeros(self.scene.n_tri, np.int32)
        _lib.check(_lib.lib().ams_scene_plane_ids(self.handle, out.ctypes.data_as(C.POINTER(C.c_int32))),
                   "ams_scene_plane_ids")
        return out

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().ams_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _array(spec: ArraySpec, keep: list) -> _lib.AmsArray:
    pos = np.ascontiguousarray(spec.pos, dtype=np.float32)
    keep.append(pos)
    a = _lib.AmsArray()
    a.n0, a.n1 = int(pos.shape[0]), int(pos.shape[1])
    a.pos = pos.ctypes.data_as(C.POINTER(C.c_float))
    d = np.asarray(spec.dir, dtype=np.float32).reshape(3)
    a.dir[0], a.dir[1], a.dir[2] = float(d[0]), float(d[1]), float(d[2])
    a.bptype = int(spec.bptype)
    a.radius = float(spec.radius)
    return a


def trace(ts: TraceScene, tx: ArraySpec, rx: ArraySpec, wave_len: float, max_reflection: int, fib_n: int,
          per_element: bool = True, gridsize: int = 5, flags: int = 0, result: Optional[torch.Tensor] = None,
          keys_cap: int = 0, tasks_per_unit: int = 0, seed_keys=None, device: str = "cuda",
          reference_launch: Optional[tuple] = None, task_info_path: Optional[str] = None):
    """Trace Tx -> Rx and accumulate into result (f32 [n_tx, n_rx, 2] on the
    GPU, zero-initialised if None).  Returns (result, stats dict, keys or None).

    reference_launch = (launch_x, launch_y, samples_per_launch, reserve_size)
    selects the reference's own discovery (optixPathTracer.cu:142-190: TEA/LCG
    random rays per launch index, per-tile min-hit gating, reserve cap)
    instead of the Fibonacci ray set; fib_n is then ignored.

    task_info_path writes one record per backward chain of every expansion
    round to that file, as the reference's path.dat (EXPORT_TASKINFO,
    optixPathTracer.cpp:244-256); read_task_info() loads it."""
    if result is None:
        result = torch.zeros(tx.n, rx.n, 2, dtype=torch.float32, device=device)
    if result.dtype != torch.float32 or not result.is_contiguous() or not result.is_cuda or \
            tuple(result.shape) != (tx.n, rx.n, 2):
        raise ValueError("result must be a contiguous f32 CUDA tensor [n_tx, n_rx, 2]")
    keep: list = []
    a_tx, a_rx = _array(tx, keep), _array(rx, keep)
    o = _lib.AmsTraceOpts()
    o.wave_len = float(wave_len)
    o.max_reflection = int(max_reflection)
    o.gridsize = int(gridsize)
    o.fib_n = int(fib_n)
    o.per_element = 1 if per_element else 0
    o.flags = int(flags)
    o.tasks_per_unit = int(tasks_per_unit)
    if reference_launch is not None:
        lx, ly, spl, res = (int(v) for v in reference_launch)
        o.discovery_mode = 1
        o.launch_x, o.launch_y, o.samples_per_launch, o.reserve_size = lx, ly, spl, res
    o.stream = C.c_void_p(torch.cuda.current_stream(result.device).cuda_stream)
    if task_info_path is not None:
        o.task_info_path = os.fsencode(task_info_path)
    keys = None
    if keys_cap > 0:
        keys = np.zeros((keys_cap, 2), dtype=np.uint64)
        o.keys_out = keys.ctypes.data_as(C.POINTER(C.c_uint64))
        o.keys_cap = int(keys_cap)
    if seed_keys is not None:
        seed = np.ascontiguousarray(seed_keys, dtype=np.uint64).reshape(-1, 2)
        keep.append(seed)
        o.seed_keys = seed.ctypes.data_as(C.POINTER(C.c_uint64))
        o.n_seed = int(seed.shape[0])
    st = _lib.AmsTraceStats()
    _lib.check(_lib.lib().ams_trace_channel(ts.handle, C.byref(a_tx), C.byref(a_rx), C.byref(o),
                                            C.c_void_p(result.data_ptr()), C.byref(st)), "ams_trace_channel")
    stats = st.as_dict()
    if keys is not None:
        keys = keys[: min(keys_cap, stats["records_unique"])]
    return result, stats, keys


# ams_task_info / ExportedTaskInfo (optixPathTracer.h:40-52): 52 bytes
TASK_INFO_DTYPE = np.dtype([("successful", np.uint8), ("pad_", np.uint8, 3), ("launch_angle", np.float32),
                            ("receive_angle", np.float32), ("loss", np.float32), ("phase", np.float32),
                            ("src", np.int32), ("dst", np.int32), ("hit_cnt", np.int32),
                            ("hit_loss_idx", np.int32, 5)])
assert TASK_INFO_DTYPE.itemsize == 52


def read_task_info(path: str) -> np.ndarray:
    """Records written by trace(task_info_path=...), in file order."""
    return np.fromfile(path, dtype=TASK_INFO_DTYPE)


_job_streams: dict = {}


def trace_many(jobs: list, device: str = "cuda", max_concurrent: Optional[int] = None) -> list:
    """Trace independent channel sets concurrently.  Each job is a dict of
    trace()'s arguments (ts, tx, rx, wave_len, max_reflection, fib_n, and any
    keywords); at most max_concurrent (default: all) run at once, each from
    its own host thread on its own stream, so the host steps of one trace
    (record sort, path list, expansion rounds) and the tails of its kernels
    overlap the kernels of the others.  Every job needs its own TraceScene
    handle: a handle's scratch serves one call at a time.  Results are those
    of calling trace() on each job in turn; the caller's current stream is
    ordered after all of them on return."""
    import threading

    if len({id(j["ts"]) for j in jobs}) != len(jobs):
        raise ValueError("trace_many: every job needs its own TraceScene handle")
    dev = torch.device(device)
    if dev.index is None:
        dev = torch.device(dev.type, torch.cuda.current_device())
    workers = max(1, min(len(jobs), max_concurrent or len(jobs)))
    caller = torch.cuda.current_stream(dev)
    streams = _job_streams.setdefault(dev.index, [])
    while len(streams) < workers:
        streams.append(torch.cuda.Stream(dev))
    jobs = [dict(j) for j in jobs]
    for j in jobs:  # results zero-filled on the caller's stream come first
        if j.get("result") is None:
            j["result"] = torch.zeros(j["tx"].n, j["rx"].n, 2, dtype=torch.float32, device=dev)
    out: list = [None] * len(jobs)
    err: list = [None] * len(jobs)
    lock = threading.Lock()
    nxt = [0]

    def run(w):
        s = streams[w]
        s.wait_stream(caller)
        with torch.cuda.device(dev), torch.cuda.stream(s):
            while True:
                with lock:
                    k = nxt[0]
                    nxt[0] += 1
                if k >= len(jobs):
                    return
                try:
                    jobs[k]["result"].record_stream(s)
                    out[k] = trace(device=str(dev), **jobs[k])
                except BaseException as e:  # re-raised in the caller's thread
                    err[k] = e

    threads = [threading.Thread(target=run, args=(w,)) for w in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for e in err:
        if e is not None:
            raise e
    for w in range(workers):
        caller.wait_stream(streams[w])
    return out


def trace_workload(w: Workload, ts: Optional[TraceScene] = None, **kw):
    ts = ts or TraceScene(w.scene)
    res, stats, keys = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                             w.gridsize, **kw)
    return torch.complex(res[..., 0], res[..., 1]), stats, keys


def coverage_db(H: torch.Tensor, floor_db: float = -150.0) -> torch.Tensor:
    """Coverage map 20 log10 |sum_tx H| with the floor of go.py:849-851, :944-964."""
    amp = H.sum(0).abs()
    db = 20 * torch.log10(amp)
    db = torch.where(torch.isnan(db), torch.full_like(db, floor_db), db)
    return db.clamp(min=floor_db)
