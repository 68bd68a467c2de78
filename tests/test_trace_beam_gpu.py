"""Backward queries answered from beam candidate lists (k_beam) against the
hierarchy walk.

For a (path, Rx) pair whose interactions are all reflections, k_beam lists
per leg every triangle within a margin of the convex beam the leg sweeps
over all Tx elements; a lane's closest hit among those candidates, under
the walk's rules (the query bound, the skipped plane, ties to the smaller
original index), is the walk's closest hit.  So the two must agree exactly:
the same record sets, rounds, tasks, chains and successes, and bit-identical
H (H is accumulated in fixed point, independent of order).  Cases: the
room with its 16-element AP (lists forced on, flags bit 15), the
metasurface channel F (on by default: 1024 Tx elements), the reference's
ray set, the door plate and the 4321 scene with a door (paths through a
penetrable surface walk every leg), the office at reduced sizes, and the
office's device-built rounds.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OFF, FORCE = 16384, 32768


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _keyset(k):
    return set(map(tuple, np.asarray(k, dtype=np.uint64).tolist()))


def _run(ts, w, flags, fib=None, **kw):
    import torch

    from automs_amd.tracer import trace
    res, st, keys = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n if fib is None else fib,
                          w.per_element, w.gridsize, flags=flags, keys_cap=1 << 20, **kw)
    torch.cuda.synchronize()
    return res.double().cpu().numpy(), _keyset(keys), st


def _same(beam, walk, label, expect_lists=True):
    (Ha, ka, sa), (Hb, kb, sb) = beam, walk
    _flush(f"  {label}: keys {len(ka)} vs {len(kb)}, tasks {sa['tasks']} vs {sb['tasks']}, successes "
           f"{sa['successes']} vs {sb['successes']}; backward queries {sa['bwd_queries']}, walked "
           f"{sa['bwd_walks']} vs {sb['bwd_walks']}; lists for {sa['beam_pairs']} pairs, "
           f"{sa['beam_cands']} candidates, {sa['beam_walk_legs']} legs left to the walk; nodes "
           f"{sa['bwd_nodes'] + sa['beam_nodes']} vs {sb['bwd_nodes']}")
    assert ka == kb
    for f in ("rounds", "tasks", "chains", "successes", "bwd_queries"):
        assert sa[f] == sb[f], f
    assert sb["bwd_walks"] == sb["bwd_queries"] and sb["beam_pairs"] == 0
    assert np.array_equal(Ha, Hb)
    if expect_lists:
        assert sa["beam_pairs"] > 0 and sa["bwd_walks"] < sa["bwd_queries"]


@pytest.fixture(scope="module")
def room(gpu):
    from automs_amd import scenes
    from automs_amd.tracer import TraceScene
    w = scenes.room_boxes(fib_n=1 << 12)
    ts = TraceScene(w.scene)
    yield w, ts
    ts.close()


@pytest.mark.parametrize("extra", [0, 16, 8192])
def test_room_lists_forced(room, extra):
    """16 Tx elements (lists forced on); with the scene in LDS, from global
    memory (bit 4) and with the device-built rounds (bit 13)."""
    w, ts = room
    _flush(f"case: beam lists, room_boxes, forced, extra flags {extra}")
    _same(_run(ts, w, FORCE | extra), _run(ts, w, OFF | extra), "room")


def test_room_reference_rays(room):
    w, ts = room
    _flush("case: beam lists, room_boxes, reference rays, forced")
    kw = dict(reference_launch=(64, 64, 1, 8))
    _same(_run(ts, w, FORCE, fib=0, **kw), _run(ts, w, OFF, fib=0, **kw), "room, reference rays")


def test_metasurface_channel_default(gpu):
    """F of config 1's room (1024 metasurface elements): lists by default,
    read with wave-uniform scalar loads (every wave's lanes share one Rx) and
    per lane (flags bit 16)."""
    from automs_amd import scenes
    from automs_amd.tracer import TraceScene
    w = scenes.room_boxes_ms(fib_n=1 << 12, ms_fib_n=1 << 16)["F"]
    _flush(f"case: beam lists, room_boxes_ms F, {w.tx.n} Tx x {w.rx.n} Rx, default")
    ts = TraceScene(w.scene)
    beam, lane, walk = _run(ts, w, 0), _run(ts, w, 65536), _run(ts, w, OFF)
    ts.close()
    _same(beam, walk, "room_boxes_ms F")
    _same(lane, walk, "room_boxes_ms F, per-lane list reads")


def test_door_plate(gpu):
    """Paths through the penetrable door walk every leg; reflections use the
    lists."""
    import _door_scene as P

    from automs_amd.scenes import ArraySpec, Scene, Workload
    from automs_amd.tracer import TraceScene
    tri = np.asarray(P.DOOR, np.float32).reshape(-1, 9)
    sc = Scene(tri=tri, mat=np.zeros(2, np.int32), mat_prop=np.asarray(P.DOOR_TABLE, np.float32).reshape(1, 19),
               refl_phase=None, names=["door"])
    tx = ArraySpec(np.asarray(P.TX, np.float32).reshape(1, 1, 3), np.array([1, 0, 0], np.float32), 0, 0.5e-5)
    rx = ArraySpec(P.door_rx(), np.zeros(3, np.float32), 0, P.RX_R)
    w = Workload("door", sc, tx, rx, P.WL, 1, 1 << 16, True, gridsize=1)
    _flush("case: beam lists, door plate, forced")
    ts = TraceScene(w.scene)
    beam, walk = _run(ts, w, FORCE), _run(ts, w, OFF)
    ts.close()
    _same(beam, walk, "door plate", expect_lists=False)


def test_scene_4321_with_door(gpu):
    """The 4321 scene with a door leaf (penetrable, -0.7): 36 AP elements,
    lists forced on."""
    from automs_amd import channel, geometry
    from automs_amd.scenes import ArraySpec, Workload
    from automs_amd.tracer import TraceScene
    import tempfile
    from pathlib import Path
    wl = 3e8 / 60e9
    with tempfile.TemporaryDirectory() as d:
        (Path(d) / "door.txt").write_text("-4.0 3.0 0.6\n-4.0 3.0 -1.4\n-4.0 4.0 -1.4\n0\n")
        doors = channel.get_doors(d, 0.3)
    sc = channel.load_scene_folder("tests/golden/scene_4321", door_buffer=doors)
    src = geometry.gen_pos_matrix([-3, 3.4, -0.51], [-1, 0, 0], [0, 0, -1], (6, 6), wl / 2).numpy()
    dst = geometry.gen_pos_from_endpoint([-6, 2.4, 42], [-1, 6.5, 40], -0.51).numpy()
    w = Workload("4321+door", sc, ArraySpec(src, np.array([-1, 0, 0], np.float32), 2, 0.5e-5),
                 ArraySpec(dst, np.zeros(3, np.float32), 0, 0.06), wl, 4, 1 << 13, True)
    _flush(f"case: beam lists, 4321 + door, {w.tx.n} Tx x {w.rx.n} Rx, forced")
    ts = TraceScene(w.scene)
    beam, walk = _run(ts, w, FORCE), _run(ts, w, OFF)
    ts.close()
    _same(beam, walk, "4321 + door")


@pytest.mark.parametrize("name", ["H_d", "F0"])
def test_office_reduced(gpu, name):
    """The office (100,046 triangles, 5 reflections) at reduced sizes: H_d
    with a 4 x 4 AP (forced), F0 from a 48 x 48 metasurface (default)."""
    from automs_amd import scenes
    from automs_amd.tracer import TraceScene
    w = scenes.office(fib_n=1 << 16, ms_fib_n=1 << 16, n_rx=(40, 28), ap=(4, 4), max_reflection=5)[name]
    _flush(f"case: beam lists, office {name}, {w.tx.n} Tx x {w.rx.n} Rx")
    ts = TraceScene(w.scene)
    flags = FORCE if name == "H_d" else 0
    beam, walk = _run(ts, w, flags), _run(ts, w, OFF)
    ts.close()
    _same(beam, walk, f"office {name}")
    if name == "F0":
        # 2304 Tx: every wave shares one Rx, and the legs whose lists overflow
        # walk the hierarchy as one wave (scalar loads); H above is bitwise the
        # per-lane walk's
        assert beam[2]["beam_walk_legs"] > 0 and beam[2]["bwd_walks"] > 0


def test_office_device_rounds(gpu):
    """Office H_d on its full 256 x 256 grid (8 x 8 AP, 2^18 rays): device-built
    rounds, lists forced on (64 elements)."""
    from automs_amd import scenes
    from automs_amd.tracer import TraceScene
    w = scenes.office(fib_n=1 << 18)["H_d"]
    _flush(f"case: beam lists, office H_d device rounds, {w.tx.n} Tx x {w.rx.n} Rx")
    ts = TraceScene(w.scene)
    beam, walk = _run(ts, w, FORCE), _run(ts, w, OFF)
    ts.close()
    _same(beam, walk, "office H_d, device rounds")
    _flush(f"  evaluation {beam[2]['t_evaluate_ms']:.1f} vs {walk[2]['t_evaluate_ms']:.1f} ms")
