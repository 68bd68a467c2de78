"""ctypes wrapper of the FP64 CPU oracle (oracle/ams_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the automs_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libams_oracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp, fp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int), \
            C.POINTER(C.c_long)
        u64p = C.POINTER(C.c_uint64)
        L.or_scene_create.argtypes = [C.c_int, fp, ip, C.c_int, fp, fp]
        L.or_scene_create.restype = vp
        L.or_scene_destroy.argtypes = [vp]
        L.or_scene_plane_ids.argtypes = [vp, ip]
        L.or_closest_hit.argtypes = [vp, dp, dp, C.c_double, C.c_int, dp]
        L.or_closest_hit.restype = C.c_int
        L.or_beam_pattern.argtypes = [dp, dp, C.c_int]
        L.or_beam_pattern.restype = C.c_double
        L.or_fib_dir.argtypes = [C.c_long, C.c_long, dp]
        L.or_set_tie_eps.argtypes = [C.c_double]
        L.or_set_skip_ties.argtypes = [C.c_int]
        L.or_set_tie_mask.argtypes = [C.POINTER(C.c_ubyte), C.c_long]
        L.or_discover.argtypes = [vp, fp, C.c_int, fp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_long,
                                  C.c_int, C.c_long, C.c_long, u64p, C.c_long, lp]
        L.or_discover.restype = C.c_long
        L.or_trace_ray_log.argtypes = [vp, dp, dp, fp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, ip,
                                       C.c_int, ip]
        L.or_trace_ray_log.restype = C.c_int
        L.or_evaluate.argtypes = [vp, fp, C.c_int, fp, C.c_int, C.c_int, C.c_double, C.c_double, fp, fp, C.c_int,
                                  C.c_int, C.c_int, u64p, C.c_int, ip, ip, C.c_long, C.c_int, dp,
                                  C.POINTER(C.c_ubyte), lp]
        L.or_backward_one.argtypes = [vp, dp, dp, C.c_int, C.c_int, C.POINTER(C.c_uint), C.c_double, C.c_double,
                                      dp, dp, C.c_int, C.c_int, dp]
        L.or_backward_one.restype = C.c_int
        L.or_trace_channel.argtypes = [vp, fp, C.c_int, C.c_int, fp, C.c_int, C.c_int, C.c_double, C.c_int,
                                       C.c_double, C.c_double, fp, fp, C.c_int, C.c_int, C.c_int, C.c_long, C.c_int,
                                       C.c_int, u64p, C.c_long, dp, lp, u64p, C.c_long, lp]
        L.or_trace_channel.restype = C.c_int
        L.or_discover_ref.argtypes = [vp, fp, C.c_int, fp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_int, u64p, C.c_long, lp]
        L.or_tea.argtypes = [C.c_uint, C.c_uint, C.POINTER(C.c_uint), C.c_int, C.POINTER(C.c_uint)]
        L.or_tea.restype = None
        L.or_lcg.argtypes = [C.c_uint]
        L.or_lcg.restype = C.c_uint
        L.or_ref_rays.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), fp]
        L.or_ref_rays.restype = None
        L.or_discover_ref.restype = C.c_long
        L.or_discover_ref_log.argtypes = [vp, fp, C.c_int, fp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_long, C.c_long, ip,
                                          C.c_long]
        L.or_discover_ref_log.restype = C.c_long
        _lib = L
    return _lib


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n: int) -> None:
    os.environ["OMP_NUM_THREADS"] = str(n)


class OracleScene:
    def __init__(self, tri, mat, mat_prop, refl_phase=None):
        L = lib()
        self._tri, tp = _f(np.asarray(tri).reshape(-1))
        self._mat = np.ascontiguousarray(mat, dtype=np.int32)
        self._mp, mpp = _f(np.asarray(mat_prop).reshape(-1))
        self._rp = None
        rpp = None
        if refl_phase is not None:
            self._rp, rpp = _f(refl_phase)
        self.n_tri = len(self._mat)
        self.h = L.or_scene_create(self.n_tri, tp, self._mat.ctypes.data_as(C.POINTER(C.c_int)),
                                   int(np.asarray(mat_prop).reshape(-1, 19).shape[0]), mpp, rpp)

    @classmethod
    def from_scene(cls, s):
        return cls(s.tri, s.mat, s.mat_prop, s.refl_phase)

    def plane_ids(self):
        out = np.zeros(self.n_tri, np.int32)
        lib().or_scene_plane_ids(self.h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def closest_hit(self, o, d, tmin=1e-5, use_bvh=True):
        _, op = _d(o)
        _, dp = _d(d)
        t = C.c_double()
        i = lib().or_closest_hit(self.h, op, dp, tmin, 1 if use_bvh else 0, C.byref(t))
        return i, t.value

    def close(self):
        if self.h:
            lib().or_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def beam_pattern(d1, d2, t):
    _, a = _d(d1)
    _, b = _d(d2)
    return lib().or_beam_pattern(a, b, int(t))


def fib_dir(k, n):
    out = np.zeros(3)
    lib().or_fib_dir(int(k), int(n), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def set_tie_eps(e):
    lib().or_set_tie_eps(float(e))


class tie_bookkeeping:
    """Context for tie-attributed parity checks: inside it, discovery drops
    the records of rays (Fibonacci mode) or launch threads (reference mode)
    that meet a near tie when skip=True, and evaluation marks every (Tx, Rx)
    entry with a near-tie chain in .mask ([n_src, n_dst] bool) when n_src,
    n_dst are given.  eps: the tie tolerance in metres."""

    def __init__(self, eps=1e-6, skip=False, n_src=0, n_dst=0):
        self.eps, self.skip = eps, skip
        self._m = np.zeros(max(1, n_src * n_dst), np.uint8)
        self.shape = (n_src, n_dst)
        self.mask = None

    def __enter__(self):
        L = lib()
        L.or_set_tie_eps(float(self.eps))
        L.or_set_skip_ties(1 if self.skip else 0)
        if self.shape[0] * self.shape[1]:
            L.or_set_tie_mask(self._m.ctypes.data_as(C.POINTER(C.c_ubyte)), self._m.size)
        return self

    def __exit__(self, *exc):
        L = lib()
        L.or_set_tie_eps(1e-6)
        L.or_set_skip_ties(0)
        L.or_set_tie_mask(None, 0)
        if self.shape[0] * self.shape[1]:
            self.mask = self._m.reshape(self.shape).astype(bool)
        return False


def discover(sc: OracleScene, src, dst, dst_r, max_reflection, gridsize, fib_n, per_element, ray_begin=0,
             ray_end=None, cap=1 << 24):
    src_a, sp = _f(np.asarray(src).reshape(-1, 3))
    dst = np.asarray(dst)
    n0, n1 = dst.shape[0], dst.shape[1]
    dst_a, dp = _f(dst.reshape(-1, 3))
    n_src = src_a.shape[0]
    if ray_end is None:
        ray_end = n_src * fib_n if per_element else fib_n
    out = np.zeros((cap, 2), np.uint64)
    st = np.zeros(4, np.int64)
    m = lib().or_discover(sc.h, sp, n_src, dp, n0, n1, dst_r, max_reflection, gridsize, fib_n,
                          1 if per_element else 0, ray_begin, ray_end, out.ctypes.data_as(C.POINTER(C.c_uint64)),
                          cap, st.ctypes.data_as(C.POINTER(C.c_long)))
    if m < 0:
        raise RuntimeError("oracle discovery cap exceeded")
    return out[:m].copy(), dict(rays=int(st[0]), queries=int(st[1]), unique=int(st[2]), ties=int(st[3]))


def tea(v0: int, v1: int, key, rounds: int):
    k = (C.c_uint * 4)(*key)
    out = (C.c_uint * 2)()
    lib().or_tea(v0, v1, k, rounds, out)
    return int(out[0]), int(out[1])


def lcg(state: int) -> int:
    return int(lib().or_lcg(state))


def ref_rays(launch_x: int, launch_y: int, samples: int, n_src: int):
    """(element [n], direction f32 [n, 3]) of every ray of a reference launch,
    ordered by launch index t = ix * launch_y + iy, then sample."""
    n = launch_x * launch_y * samples
    elem = np.zeros(n, np.int32)
    d = np.zeros((n, 3), np.float32)
    lib().or_ref_rays(launch_x, launch_y, samples, n_src, elem.ctypes.data_as(C.POINTER(C.c_int)),
                      d.ctypes.data_as(C.POINTER(C.c_float)))
    return elem, d


def discover_ref(sc: OracleScene, src, dst, dst_r, max_reflection, gridsize, launch_x, launch_y, samples,
                 reserve, cap=1 << 24, gating: bool = True):
    """Reference-compatible discovery (TEA/LCG rays, min_hc gating, reserve);
    gating False records every Rx hit of the same rays (for subset checks)."""
    src_a, sp = _f(np.asarray(src).reshape(-1, 3))
    dst = np.asarray(dst)
    n0, n1 = dst.shape[0], dst.shape[1]
    dst_a, dp = _f(dst.reshape(-1, 3))
    out = np.zeros((cap, 2), np.uint64)
    st = np.zeros(5, np.int64)
    m = lib().or_discover_ref(sc.h, sp, src_a.shape[0], dp, n0, n1, dst_r, max_reflection, gridsize, launch_x,
                              launch_y, samples, reserve, 1 if gating else 0,
                              out.ctypes.data_as(C.POINTER(C.c_uint64)), cap,
                              st.ctypes.data_as(C.POINTER(C.c_long)))
    if m < 0:
        raise RuntimeError("oracle discovery cap exceeded")
    return out[:m].copy(), dict(rays=int(st[0]), queries=int(st[1]), unique=int(st[2]), raw=int(st[3]),
                                tie_threads=int(st[4]))


def discover_ref_log(sc: OracleScene, src, dst, dst_r, max_reflection, gridsize, launch_x, launch_y, samples,
                     reserve, t_begin, t_end, gating: bool = True, cap=1 << 22):
    """The Rx events of launch threads [t_begin, t_end) of discover_ref in the
    order each thread meets them: int array [n, 5] of (thread, sample, h,
    tile, recorded)."""
    src_a, sp = _f(np.asarray(src).reshape(-1, 3))
    dst = np.asarray(dst)
    n0, n1 = dst.shape[0], dst.shape[1]
    dst_a, dp = _f(dst.reshape(-1, 3))
    log = np.zeros(cap, np.int32)
    n = lib().or_discover_ref_log(sc.h, sp, src_a.shape[0], dp, n0, n1, dst_r, max_reflection, gridsize, launch_x,
                                  launch_y, samples, reserve, 1 if gating else 0, t_begin, t_end,
                                  log.ctypes.data_as(C.POINTER(C.c_int)), cap)
    if n + 5 > cap:
        raise RuntimeError("oracle event log cap reached")
    return log[:n].reshape(-1, 5).copy()


def trace_ray_log(sc: OracleScene, o, d, dst, dst_r, max_reflection, gridsize, cap=256):
    dst = np.asarray(dst)
    n1 = dst.shape[1]
    dst_a, dp = _f(dst.reshape(-1, 3))
    _, op = _d(o)
    _, dd = _d(d)
    log = np.zeros(cap, np.int32)
    tie = C.c_int()
    n = lib().or_trace_ray_log(sc.h, op, dd, dp, dst_a.shape[0], dst_r, max_reflection, n1, gridsize,
                               log.ctypes.data_as(C.POINTER(C.c_int)), cap, C.byref(tie))
    return [(int(log[i]), int(log[i + 1])) for i in range(0, n, 2)], tie.value


def backward_one(sc, si, rj, h, bmask, planes, wave_len, src_r, src_dir, dst_dir, src_bp, dst_bp):
    pl = (C.c_uint * 5)(*([int(p) for p in planes] + [0xFFFFF] * (5 - len(planes))))
    _, a = _d(si)
    _, b = _d(rj)
    _, c = _d(src_dir)
    _, d = _d(dst_dir)
    out = np.zeros(3)
    ok = lib().or_backward_one(sc.h, a, b, int(h), int(bmask), pl, wave_len, src_r, c, d, int(src_bp), int(dst_bp),
                               out.ctypes.data_as(C.POINTER(C.c_double)))
    return bool(ok), complex(out[0], out[1]), float(out[2])


def trace_channel(sc: OracleScene, src, dst, wave_len, max_reflection, src_r, dst_r, src_dir, dst_dir, src_bp,
                  dst_bp, gridsize, fib_n, per_element, flags=0, seed=None, paths_cap=1 << 22):
    """Full oracle trace; returns (H complex128 [n_src, n_dst], stats, paths).
    seed: optional (path, tile) records [n, 2] u64 used instead of discovery;
    flags bit 0 = reference row wrap of the last tile column, bit 1 = no
    neighbour expansion."""
    src = np.asarray(src)
    dst = np.asarray(dst)
    s_a, sp = _f(src.reshape(-1, 3))
    d_a, dp = _f(dst.reshape(-1, 3))
    _, sdp = _f(np.asarray(src_dir).reshape(3))
    _, ddp = _f(np.asarray(dst_dir).reshape(3))
    n_src, n_dst = s_a.shape[0], d_a.shape[0]
    H = np.zeros((n_src, n_dst, 2), np.float64)
    st = np.zeros(11, np.int64)
    paths = np.zeros((paths_cap, 2), np.uint64)
    npath = C.c_long()
    seed_p, n_seed = None, 0
    if seed is not None:
        seed = np.ascontiguousarray(seed, dtype=np.uint64).reshape(-1, 2)
        seed_p, n_seed = seed.ctypes.data_as(C.POINTER(C.c_uint64)), seed.shape[0]
    rc = lib().or_trace_channel(sc.h, sp, src.shape[0], src.shape[1], dp, dst.shape[0], dst.shape[1], wave_len,
                                max_reflection, src_r, dst_r, sdp, ddp, src_bp, dst_bp, gridsize, fib_n,
                                1 if per_element else 0, flags, seed_p, n_seed,
                                H.ctypes.data_as(C.POINTER(C.c_double)),
                                st.ctypes.data_as(C.POINTER(C.c_long)), paths.ctypes.data_as(C.POINTER(C.c_uint64)),
                                paths_cap, C.byref(npath))
    if rc != 0:
        raise RuntimeError("oracle trace failed")
    names = ["rays", "fwd_queries", "records_unique", "fwd_ties", "paths", "tasks", "rounds", "chains",
             "successes", "bwd_queries", "bwd_ties"]
    return H[..., 0] + 1j * H[..., 1], dict(zip(names, map(int, st))), paths[: min(npath.value, paths_cap)]


def trace_workload(w, flags=0, seed=None, sc=None):
    sc = sc or OracleScene.from_scene(w.scene)
    return trace_channel(sc, w.tx.pos, w.rx.pos, w.wave_len, w.max_reflection, w.tx.radius, w.rx.radius, w.tx.dir,
                         w.rx.dir, w.tx.bptype, w.rx.bptype, w.gridsize, w.fib_n, w.per_element, flags, seed)


def key_fields(hi, lo):
    """Decode a 128-bit path key into (h, b, planes, tile)."""
    hi, lo = int(hi), int(lo)
    h = hi >> 61
    b = (hi >> 56) & 31
    p0 = (hi >> 36) & 0xFFFFF
    p1 = (hi >> 16) & 0xFFFFF
    p2 = ((hi & 0xFFFF) << 4) | (lo >> 60)
    p3 = (lo >> 40) & 0xFFFFF
    p4 = (lo >> 20) & 0xFFFFF
    return h, b, [p0, p1, p2, p3, p4], lo & 0xFFFFF
