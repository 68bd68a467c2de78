"""CPU tests of the host side: the C-ABI library loads and exports every
symbol include/ams.h declares; the threaded BVH the kernels traverse gives
brute-force closest hits; plane ids agree with the oracle's; geometry
helpers and seeded scenes follow their definitions; the contraction oracle's
autograd gradient matches central differences (complex128)."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest
import torch

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    from automs_amd import _lib
    header = (ROOT / "include" / "ams.h").read_text()
    declared = set(re.findall(r"\b(ams_[a-z0-9_]+)\s*\(", header))
    L = _lib.lib()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.EXPORTED) == declared
    assert L.ams_version() == 1


def _bvh(lo, hi, leaf_max=4):
    from automs_amd import _lib
    L = _lib.lib()
    n = lo.shape[0]
    nodes = np.zeros((2 * n + 1, 8), np.float32)
    order = np.zeros(max(n, 1), np.int32)
    nn = C.c_int32()
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    _lib.check(L.ams_host_bvh_build(C.c_void_p(lo.ctypes.data), C.c_void_p(hi.ctypes.data), n, leaf_max,
                                    C.c_void_p(nodes.ctypes.data), nodes.shape[0],
                                    C.c_void_p(order.ctypes.data), C.c_void_p(C.addressof(nn))), "bvh")
    nodes = nodes[: nn.value]
    ints = nodes.view(np.int32)
    return nodes[:, 0:3], ints[:, 3], nodes[:, 4:7], ints[:, 7], order


def _mt(o, d, v0, v1, v2):
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    if det == 0:
        return None
    tv = o - v0
    u = (tv @ p) / det
    if u < 0 or u > 1:
        return None
    q = np.cross(tv, e1)
    v = (d @ q) / det
    if v < 0 or u + v > 1:
        return None
    return (e2 @ q) / det


def test_threaded_bvh_closest_hit_equals_brute_force():
    rng = np.random.default_rng(1)
    T = 200
    tri = (rng.uniform(-4, 4, (T, 1, 3)) + rng.normal(0, 0.5, (T, 3, 3)))
    lo, hi = tri.min(1), tri.max(1)
    blo, miss, bhi, info, order = _bvh(lo, hi)
    n_nodes = len(miss)
    # structure: preorder, every miss link points past its subtree, leaves cover all prims once
    assert miss[0] == n_nodes and (miss > np.arange(n_nodes)).all()
    seen = []
    for i in range(n_nodes):
        if info[i]:
            first, cnt = info[i] >> 3, info[i] & 7
            assert 1 <= cnt <= 4
            seen.extend(order[first:first + cnt])
            # leaf bounds contain their triangles
            for p in order[first:first + cnt]:
                assert (blo[i] <= lo[p] + 1e-6).all() and (bhi[i] >= hi[p] - 1e-6).all()
    assert sorted(seen) == list(range(T))
    for _ in range(200):
        o = rng.uniform(-5, 5, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        inv = 1 / np.where(d == 0, 1e-30, d)
        best, bi = np.inf, -1
        i = 0
        while i < n_nodes:  # the kernels' stackless loop
            t0 = np.minimum((blo[i] - o) * inv, (bhi[i] - o) * inv).max()
            t1 = np.maximum((blo[i] - o) * inv, (bhi[i] - o) * inv).min()
            if max(t0, 1e-5) <= min(t1, best) * (1 + 4e-7):
                if info[i] == 0:
                    i += 1
                    continue
                for p in order[(info[i] >> 3):(info[i] >> 3) + (info[i] & 7)]:
                    t = _mt(o, d, *tri[p])
                    if t is not None and 1e-5 <= t < best:
                        best, bi = t, p
            i = miss[i]
        bbest, bbi = np.inf, -1
        for p in range(T):
            t = _mt(o, d, *tri[p])
            if t is not None and 1e-5 <= t < bbest:
                bbest, bbi = t, p
        assert bi == bbi


def test_wave_walk_equals_lane_walks():
    """closest_tri_wave's rule on the threaded hierarchy: the wave enters a
    node when any lane's box test passes and a lane tests a leaf's triangles
    only when its own test passes.  Each lane's closest hit equals its own
    stackless walk's and brute force, for coherent waves (one origin, nearly
    one direction) and incoherent ones."""
    rng = np.random.default_rng(7)
    T = 160
    tri = (rng.uniform(-4, 4, (T, 1, 3)) + rng.normal(0, 0.5, (T, 3, 3)))
    blo, miss, bhi, info, order = _bvh(tri.min(1), tri.max(1))
    n_nodes = len(miss)

    def box(i, o, inv, best):
        t0 = np.minimum((blo[i] - o) * inv, (bhi[i] - o) * inv).max()
        t1 = np.maximum((blo[i] - o) * inv, (bhi[i] - o) * inv).min()
        return max(t0, 1e-5) <= min(t1, best) * (1 + 4e-7)

    def leaf(i, o, d, best, bi):
        for p in order[(info[i] >> 3):(info[i] >> 3) + (info[i] & 7)]:
            t = _mt(o, d, *tri[p])
            if t is not None and 1e-5 <= t < best:
                best, bi = t, p
        return best, bi

    visits = [0, 0]
    for coherent in (True, False):
        for _ in range(3):
            o0 = rng.uniform(-5, 5, 3)
            d0 = rng.normal(size=3)
            if coherent:
                O = o0 + rng.normal(0, 0.05, (64, 3))
                D = d0 / np.linalg.norm(d0) + rng.normal(0, 0.05, (64, 3))
            else:
                O = rng.uniform(-5, 5, (64, 3))
                D = rng.normal(size=(64, 3))
            D /= np.linalg.norm(D, axis=1, keepdims=True)
            inv = 1 / np.where(D == 0, 1e-30, D)
            best, bi = np.full(64, np.inf), np.full(64, -1)
            i = 0
            while i < n_nodes:  # one node index for the wave
                visits[coherent] += 1
                h = [box(i, O[k], inv[k], best[k]) for k in range(64)]
                if any(h):
                    if info[i] == 0:
                        i += 1
                        continue
                    for k in range(64):
                        if h[k]:
                            best[k], bi[k] = leaf(i, O[k], D[k], best[k], bi[k])
                i = miss[i]
            for k in range(64):
                lb, lbi, j = np.inf, -1, 0
                while j < n_nodes:  # the lane's own walk
                    if box(j, O[k], inv[k], lb):
                        if info[j] == 0:
                            j += 1
                            continue
                        lb, lbi = leaf(j, O[k], D[k], lb, lbi)
                    j = miss[j]
                hits = [(_mt(O[k], D[k], *tri[p]), p) for p in range(T)]
                hits = [(t, p) for t, p in hits if t is not None and t >= 1e-5]
                assert bi[k] == lbi == (min(hits)[1] if hits else -1)
                assert best[k] == lb
    assert visits[True] < visits[False]  # coherent waves visit fewer nodes


def _bvh4(lo, hi, leaf_max=4):
    from automs_amd import _lib
    L = _lib.lib()
    n = lo.shape[0]
    nodes = np.zeros((max(n, 1) + 1, 32), np.float32)
    order = np.zeros(max(n, 1), np.int32)
    nn = C.c_int32()
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    _lib.check(L.ams_host_bvh4_build(C.c_void_p(lo.ctypes.data), C.c_void_p(hi.ctypes.data), n, leaf_max,
                                     C.c_void_p(nodes.ctypes.data), nodes.shape[0],
                                     C.c_void_p(order.ctypes.data), C.c_void_p(C.addressof(nn))), "bvh4")
    nodes = nodes[: nn.value]
    blo = nodes[:, 0:12].reshape(-1, 3, 4).transpose(0, 2, 1)  # [node][child][axis]
    bhi = nodes[:, 12:24].reshape(-1, 3, 4).transpose(0, 2, 1)
    child = nodes.view(np.int32)[:, 24:28]
    return blo, bhi, child, order


EMPTY = -(1 << 31)


def _walk4(blo, bhi, child, order, tri, o, d, stack=12):
    """The kernels' four-wide walk: children sorted by entry distance, the
    nearest continued, the others pushed; leaves tested when reached."""
    inv = 1 / np.where(d == 0, 1e-30, d)
    best, bi, visits = np.inf, -1, 0
    stk, cur, depth = [], 0, 0
    while True:
        nxt = None
        if cur >= 0:
            visits += 1
            ent = []
            for c in range(4):
                if child[cur, c] == EMPTY:
                    continue
                t0 = np.minimum((blo[cur, c] - o) * inv, (bhi[cur, c] - o) * inv).max()
                t1 = np.maximum((blo[cur, c] - o) * inv, (bhi[cur, c] - o) * inv).min()
                t0 = max(t0, 1e-5)
                if t0 <= min(t1, best) * (1 + 4e-7):
                    ent.append((t0, child[cur, c]))
            ent.sort(key=lambda e: e[0])
            for _, c in reversed(ent[1:]):
                stk.append(c)
            depth = max(depth, len(stk))
            if ent:
                nxt = ent[0][1]
        else:
            code = -1 - cur
            for p in order[(code >> 3):(code >> 3) + (code & 7)]:
                t = _mt(o, d, *tri[p])
                if t is not None and 1e-5 <= t < best:
                    best, bi = t, p
        if nxt is not None:
            cur = nxt
        elif stk:
            cur = stk.pop()
        else:
            break
    assert depth <= stack
    return bi, visits


@pytest.mark.parametrize("T", [0, 1, 3, 200, 1500])
def test_four_wide_bvh_structure_and_closest_hit(T):
    rng = np.random.default_rng(7 + T)
    tri = (rng.uniform(-4, 4, (T, 1, 3)) + rng.normal(0, 0.5, (T, 3, 3)))
    lo, hi = (tri.min(1), tri.max(1)) if T else (np.zeros((0, 3)), np.zeros((0, 3)))
    blo, bhi, child, order = _bvh4(lo, hi)
    n4 = child.shape[0]
    assert n4 >= 1
    # a tree rooted at 0: every node but the root referenced exactly once
    inner = child[child >= 0]
    assert ((inner > 0) & (inner < n4)).all()
    assert sorted(inner.tolist()) == list(range(1, n4))
    # leaves cover every primitive exactly once, within their boxes; child
    # boxes of an inner child lie inside that child's box
    seen = []
    for i in range(n4):
        for c in range(4):
            r = child[i, c]
            if r == EMPTY:
                continue
            if r < 0:
                code = -1 - r
                first, cnt = code >> 3, code & 7
                assert 1 <= cnt <= 4
                for p in order[first:first + cnt]:
                    seen.append(p)
                    assert (blo[i, c] <= lo[p] + 1e-6).all() and (bhi[i, c] >= hi[p] - 1e-6).all()
            else:
                for c2 in range(4):
                    if child[r, c2] != EMPTY:
                        assert (blo[i, c] <= blo[r, c2]).all() and (bhi[i, c] >= bhi[r, c2]).all()
    assert sorted(seen) == list(range(T))
    if T == 0:
        assert (child == EMPTY).all()
        return
    for _ in range(100):
        o = rng.uniform(-5, 5, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        bi, _ = _walk4(blo, bhi, child, order, tri, o, d)
        bbest, bbi = np.inf, -1
        for p in range(T):
            t = _mt(o, d, *tri[p])
            if t is not None and 1e-5 <= t < bbest:
                bbest, bbi = t, p
        assert bi == bbi


def test_plane_ids_match_oracle():
    from automs_amd import _lib, scenes
    from oracle import oracle as O
    sc = scenes.room_boxes_scene(seed=3)
    tri = np.ascontiguousarray(sc.tri, np.float32)
    out = np.zeros(sc.n_tri, np.int32)
    _lib.check(_lib.lib().ams_host_plane_ids(C.c_void_p(tri.ctypes.data), sc.n_tri, C.c_void_p(out.ctypes.data)),
               "plane ids")
    assert np.array_equal(out, O.OracleScene.from_scene(sc).plane_ids())
    # the two triangles of every face share a plane; the room floor and the
    # box bottoms (resting on it) share one plane too
    assert (out[0::2] == out[1::2]).all()
    floor = out[2]
    assert all(out[12 * (b + 1) + 2] == floor for b in range(20))


def test_geometry_helpers():
    from automs_amd import geometry as Gm
    wl = 0.005
    pos = Gm.gen_pos_matrix([1.0, 2.0, 3.0], [1, 0, 0], [0, 0, -1], (4, 4), wl / 2)
    assert pos.shape == (4, 4, 3)
    assert torch.allclose(pos.reshape(-1, 3).mean(0), torch.tensor([1.0, 2.0, 3.0]))
    assert torch.allclose((pos[1, 0] - pos[0, 0]).norm(), torch.tensor(wl / 2), atol=1e-6)
    assert torch.allclose(pos[1, 0] - pos[0, 0], torch.tensor([0, 0, -wl / 2]), atol=1e-6)  # i axis = down
    assert torch.allclose(pos[0, 1] - pos[0, 0], torch.tensor([0, wl / 2, 0]), atol=1e-6)  # j = k x i
    ms, bh, n = Gm.gen_ms([10.0, 4.0, 1.5], [-1, 0, 0], [0, 0, -1], (32, 32), wl / 2)
    assert torch.allclose(n, torch.tensor([-1.0, 0, 0]))
    assert torch.allclose(ms.reshape(-1, 3)[:, 0], torch.tensor(9.999))
    bh = np.array(bh)
    assert bh.shape == (2, 3, 3) and np.allclose(bh[..., 0], 9.999)
    assert np.isclose(bh[..., 1].max() - bh[..., 1].min(), 32 * wl / 2, atol=1e-6)
    grid = Gm.gen_pos_from_endpoint([0.1, 9.9, 64], [0.1, 7.9, 64], 1.2)
    assert grid.shape == (64, 64, 3) and torch.allclose(grid[..., 2], torch.tensor(1.2))
    v, k, i = Gm.gen_sth_from_wall([[0, 0, 3], [0, 0, 0], [0, 4, 0]], (0.5, 0.25))
    assert torch.allclose(v, torch.tensor([0.0, 1.0, 1.5])) and torch.allclose(k.abs(), torch.tensor([1.0, 0, 0]))
    d = Gm.gen_direction_from_angle(0, 90, [0, 0, 1], [1, 0, 0])
    assert torch.allclose(d, torch.tensor([0.0, 1.0, 0.0]), atol=1e-7) or \
        torch.allclose(d, torch.tensor([0.0, -1.0, 0.0]), atol=1e-7)


def test_seeded_scene_is_deterministic():
    from automs_amd import scenes
    a = scenes.room_boxes_scene(seed=0)
    b = scenes.room_boxes_scene(seed=0)
    c = scenes.room_boxes_scene(seed=1)
    assert a.n_tri == 252 and np.array_equal(a.tri, b.tri) and not np.array_equal(a.tri, c.tri)
    gam = a.mat_prop[:, 0]
    assert gam.min() >= 0.3 and gam.max() <= 0.7 and (a.mat_prop == gam[:, None]).all()
    # boxes rest on the floor, inside the room
    boxes = a.tri[12:].reshape(20, 12, 3, 3)
    assert np.allclose(boxes[..., 2].min((1, 2)), 0.0)
    assert (boxes[..., 0] >= 0).all() and (boxes[..., 0] <= 10).all()
    assert (boxes[..., 1] >= 0).all() and (boxes[..., 1] <= 8).all()


def test_contraction_oracle_gradient_matches_finite_differences():
    """Pins the oracle of the phase-map gradient (torch autograd on CPU,
    complex128) with central differences h = 1e-5 on both objectives."""
    from oracle import contract_ref as R
    g = torch.Generator().manual_seed(0)

    def crand(*s, scale):
        return (torch.randn(*s, generator=g, dtype=torch.float64)
                + 1j * torch.randn(*s, generator=g, dtype=torch.float64)) * scale
    M, K, N, Cc = 6, 20, 40, 5
    Hd, G, F = crand(M, N, scale=3e-5), crand(M, K, scale=3e-4), crand(K, N, scale=3e-4)
    A = torch.ones(Cc, M)
    ta = torch.rand(Cc, M, generator=g, dtype=torch.float64) * 2 * math.pi
    tm = torch.rand(K, generator=g, dtype=torch.float64) * 2 * math.pi
    mask = torch.tensor([False, True, False, False, False])
    for name, fn in (("logsnr", lambda Y: R.loss_logsnr(Y, 1, 1e-9)),
                     ("capacity", lambda Y: R.loss_capacity(Y, 1e-9, 0.5, mask))):
        tmv = tm.clone().requires_grad_(True)
        loss = fn(R.received(A, ta, Hd, [G], [F], [tmv], [2 * math.pi]))
        loss.backward()
        h = 1e-5
        for k in (0, 7, 13):
            p, m = tm.clone(), tm.clone()
            p[k] += h
            m[k] -= h
            fd = (fn(R.received(A, ta, Hd, [G], [F], [p], [2 * math.pi])).item()
                  - fn(R.received(A, ta, Hd, [G], [F], [m], [2 * math.pi])).item()) / (2 * h)
            assert tmv.grad[k].item() == pytest.approx(fd, rel=1e-4, abs=1e-9), (name, k)


def test_scene_folder_loader_4321():
    """The reference's 4321 scene (data fixture): 2,689 triangles; material
    from the file-name suffix; the 19-bin table; synthetic materials appended
    in the order blackhole, reflector, door (go.py:679-737)."""
    from automs_amd import channel
    sc = channel.load_scene_folder("tests/golden/scene_4321")
    assert sc.n_tri == 2689
    assert sc.names == ["screen", "wood", "carpet", "wall3", "wall", "glass"]
    assert np.bincount(sc.mat).tolist() == [780, 1740, 12, 44, 111, 2]
    assert sc.mat_prop.shape == (6, 19) and sc.mat_prop[4, 0] == pytest.approx(0.7)
    bh = [[[0, 0, 0], [1, 0, 0], [0, 1, 0]]]
    door = [[[0, 0, 1], [1, 0, 1], [0, 1, 1]]]
    sc2 = channel.load_scene_folder("tests/golden/scene_4321", blackhole_buffer=[bh], door_buffer=[door])
    assert sc2.names[-2:] == ["blackhole", "door"] and sc2.n_tri == 2691
    assert sc2.mat_prop[-2, 0] == 0.0 and sc2.mat_prop[-1, 0] == pytest.approx(-0.7)
    walls = channel.get_walls("tests/golden/scene_4321")
    assert len(walls) == 3 and walls[0][0] == [-3.9151, 0.3959, 1.6322]


def test_torch_op_schema_registered():
    import torch

    from automs_amd import channel
    channel.register_torch_op()
    op = torch.ops.pathtrace.torch_launch_pathtrace
    names = [a.name for a in op.default._schema.arguments]
    assert names == ["result", "src", "dst", "src_aabb_buffer", "dst_aabb_buffer", "scene_triangle_buffer", "mat",
                     "samples_per_launch", "launch_x", "launch_y", "wave_len", "max_reflection", "reserve_size",
                     "mat_prop", "src_dir", "dst_dir", "optix_log_level", "src_bptype", "dst_bptype",
                     "dst_girdsize", "src_r", "dst_r"]


def test_surrogate_placements_and_features():
    """Placement poses on the four walls, neighbour moves stay in range, and
    the 9 input channels follow their definitions."""
    from automs_amd import surrogate as S
    from automs_amd.scenes import SplitMix64
    poses = [S.Placement(w, 0.5, 1.5).pose() for w in range(4)]
    assert poses[0][:2] == (0.0, 4.0) and poses[1][:2] == (10.0, 4.0)
    assert poses[2][:2] == (5.0, 0.0) and poses[3][:2] == (5.0, 8.0)
    rng = SplitMix64(0)
    p = S.Placement(1, 0.5, 1.5)
    for _ in range(200):
        p = S.neighbour(p, rng)
        assert p.wall in (0, 1, 2, 3) and 0.1 <= p.s <= 0.9 and 0.6 <= p.z <= 2.4
    pos, bh, nrm = S.Placement(3, 0.25, 2.0).ms()
    assert pos.shape == (32, 32, 3) and np.allclose(nrm, [0, -1, 0], atol=1e-6)
    assert np.allclose(pos[..., 1], 8.0 - 0.001, atol=1e-5)
    occ = S.occupancy(0)
    assert occ.shape == (64, 64) and 0 < occ.max() <= 0.5 + 1e-6
    f = S.features(occ, S.Placement(1, 0.5, 1.5))
    assert f.shape == (9, 64, 64)
    assert np.allclose(f[0], occ) and np.isclose(f[7].mean(), 0.5)
    # the Rx closest to the metasurface (x = 9.9, y ~ 4) has the largest cosine to its normal
    r = f[5]
    i, j = np.unravel_index(np.argmin(r), r.shape)
    assert i == 63 and abs(j - 32) <= 1


def test_task_info_record_layout():
    """The task export record: the numpy dtype tracer.read_task_info uses has
    the field offsets of ams_task_info (include/ams.h), the 52-byte layout of
    the reference's ExportedTaskInfo (optixPathTracer.h:40-52: a bool padded
    to 4 bytes, four floats, three ints, five ints)."""
    from automs_amd.tracer import TASK_INFO_DTYPE

    class Rec(C.Structure):
        _fields_ = [("successful", C.c_uint8), ("pad_", C.c_uint8 * 3), ("launch_angle", C.c_float),
                    ("receive_angle", C.c_float), ("loss", C.c_float), ("phase", C.c_float), ("src", C.c_int32),
                    ("dst", C.c_int32), ("hit_cnt", C.c_int32), ("hit_loss_idx", C.c_int32 * 5)]
    assert C.sizeof(Rec) == TASK_INFO_DTYPE.itemsize == 52
    for name, *_ in Rec._fields_:
        assert getattr(Rec, name).offset == TASK_INFO_DTYPE.fields[name][1], name
    header = (ROOT / "include" / "ams.h").read_text()
    block = header[header.index("typedef struct {\n  uint8_t successful;"):]
    block = block[: block.index("} ams_task_info;")]
    for name, *_ in Rec._fields_:
        assert name.rstrip("_") in block, name


def test_fused_adam_flat_gradient_view():
    """FusedAdam._flat_grad returns a view of the gradients' buffer when they
    lie back to back in parameter order (as the fused operand's backward
    writes them), and their concatenation otherwise (separate tensors, the
    wrong order, a missing gradient)."""
    import torch
    from automs_amd.phasemap import FusedAdam
    a = torch.nn.Parameter(torch.zeros(3, 4))
    b = torch.nn.Parameter(torch.zeros(2, 5))
    opt = FusedAdam([a, b])
    buf = torch.arange(22, dtype=torch.float32)
    a.grad, b.grad = buf[:12].view(3, 4), buf[12:].view(2, 5)
    g = opt._flat_grad()
    assert g.data_ptr() == buf.data_ptr() and torch.equal(g, buf)
    a.grad, b.grad = buf[10:22].view(3, 4), buf[:10].view(2, 5)  # same buffer, other order
    g = opt._flat_grad()
    assert g.data_ptr() != buf.data_ptr()
    assert torch.equal(g, torch.cat([buf[10:22], buf[:10]]))
    a.grad, b.grad = torch.ones(3, 4), torch.full((2, 5), 2.0)
    assert torch.equal(opt._flat_grad(), torch.cat([torch.ones(12), torch.full((10,), 2.0)]))
    b.grad = None
    assert torch.equal(opt._flat_grad(), torch.cat([torch.ones(12), torch.zeros(10)]))


def test_contraction_oracle_fp32_mode_and_start_helpers():
    """The oracle's FP32 mode (the host baseline of the phase-map optimiser)
    returns complex64 within 2e-6 of its FP64 result; ap_amplitude follows
    go.py:59-66 (1/sqrt(M/36), the corners of a 6x6 AP off); start_phases
    is seeded and in [0, 2 pi)."""
    from automs_amd.phasemap import ap_amplitude, start_phases
    from oracle import contract_ref as R
    g = torch.Generator().manual_seed(3)

    def crand(*s, scale):
        return (torch.randn(*s, generator=g, dtype=torch.float64)
                + 1j * torch.randn(*s, generator=g, dtype=torch.float64)) * scale
    M, K, N, Cc = 16, 64, 128, 7
    Hd, G, F = crand(M, N, scale=3e-5), crand(M, K, scale=3e-4), crand(K, N, scale=3e-4)
    ta = torch.rand(Cc, M, generator=g, dtype=torch.float64) * 2 * math.pi
    tm = (torch.rand(K, generator=g, dtype=torch.float64) - 0.3) * 6 * math.pi
    A = ap_amplitude(Cc, M)
    y64 = R.received(A, ta, Hd, [G], [F], [tm], [5.0])
    y32 = R.received(A, ta, Hd, [G], [F], [tm], [5.0], dtype=torch.float32)
    assert y64.dtype == torch.complex128 and y32.dtype == torch.complex64
    assert ((y32.to(torch.complex128) - y64).norm() / y64.norm()).item() < 2e-6
    assert torch.allclose(A, torch.full((Cc, M), 1.5))
    A36 = ap_amplitude(3, 36)
    assert (A36[:, [0, 5, 30, 35]] == 0).all() and (A36.sum(1) == 32).all()
    a1, b1 = start_phases(K, M, Cc, seed=4)
    a2, b2 = start_phases(K, M, Cc, seed=4)
    assert torch.equal(a1, a2) and torch.equal(b1, b2) and a1.shape == (1, K) and b1.shape == (Cc, M)
    assert float(a1.min()) >= 0 and float(a1.max()) < 2 * math.pi


def test_fused_adam_rejects_a_wrong_best_buffer():
    """FusedAdam.step checks the best-parameter buffer (size, dtype,
    layout) before it launches anything."""
    import torch
    from automs_amd.phasemap import FusedAdam
    a = torch.nn.Parameter(torch.zeros(5))
    opt = FusedAdam([a])
    a.grad = torch.ones(5)
    loss = torch.tensor(1.0)
    for bad in (torch.zeros(4), torch.zeros(5, dtype=torch.float64), torch.zeros(10)[::2]):
        with pytest.raises(ValueError):
            opt.step(loss=loss, best=bad)


def test_split_count_rule():
    """Reduction splits of the per-wave contraction kernel: about 512 waves
    of 32 x 64 output tiles for short reductions; for long ones (200 and
    more 32-deep blocks) the most splits that keep the waves at or below
    1024, one per SIMD.  Values at the shapes they were measured on."""
    from automs_amd.phasemap import split_count

    def waves(Ra, Rb, Kd):
        return ((Ra + 31) // 32) * ((Rb + 63) // 64) * split_count(Ra, Rb, Kd)

    # the reference run's 30 x 18900 x 32036 product and its backward
    assert split_count(30, 18900, 32036) == 3 and split_count(30, 32036, 18900) == 2
    # the office three-surface backward and chain products; its forward fills the GPU unsplit
    assert split_count(30, 6976, 65536) == 9 and split_count(30, 6912, 6912) == 9
    assert split_count(30, 65536, 6976) == 1
    # config 1's short reductions keep the 512-wave target
    assert split_count(30, 4096, 1040) == 8 and split_count(30, 1040, 4096) == 30
    for shape in [(30, 18900, 32036), (30, 32036, 18900), (30, 6976, 65536), (30, 6912, 6912), (64, 9000, 20000)]:
        assert waves(*shape) <= 1024
        assert split_count(*shape) <= (shape[2] + 31) // 32
