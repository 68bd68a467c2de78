"""Penetrable materials on the GPU (the transmit branch of
optixPathTracer.cu:307, :327 and the |value| loss of :363) against the FP64
oracle and closed forms, in both discovery modes.

  * the door scene of tests/_door_scene.py: one door plate whose
    table is penetrable (< 0) below 47.5 degrees of incidence and reflecting
    above; the Fibonacci record set is the hand-enumerated one, and H of the
    round-1 tasks is the closed-form line-of-sight + door-reflection +
    door-transmission sum (FP32 bar: 1e-5 relative per Rx);
  * the reference's 4321 scene with a door leaf built by get_doors
    (go.py:554-594) and the door material -0.7 (go.py:726-733), 6 x 6 AP,
    4 reflections: records, successes and H against the oracle with the
    tie-attributed bars, for Fibonacci rays and for the reference's
    100 x 100 TEA/LCG launch.
"""
import numpy as np
import pytest

from _parity import (check_discovery_fib, check_discovery_ref, check_h, flush, keyset, oracle_trace,
                     to_complex)

pytestmark = pytest.mark.gpu

WL = 3e8 / 60e9


def _door_workload():
    import _door_scene as P

    from automs_amd.scenes import ArraySpec, Scene, Workload
    tri = np.asarray(P.DOOR, np.float32).reshape(-1, 9)
    sc = Scene(tri=tri, mat=np.zeros(2, np.int32), mat_prop=np.asarray(P.DOOR_TABLE, np.float32).reshape(1, 19),
               refl_phase=None, names=["door"])
    tx = ArraySpec(np.asarray(P.TX, np.float32).reshape(1, 1, 3), np.array([1, 0, 0], np.float32), 0, 0.5e-5)
    rx = ArraySpec(P.door_rx(), np.zeros(3, np.float32), 0, P.RX_R)
    return P, Workload("door", sc, tx, rx, WL, 1, 1 << 18, True, gridsize=1)


def test_door_plate_fibonacci_by_hand(gpu):
    import torch

    from automs_amd.tracer import TraceScene, trace
    from oracle import oracle as O
    P, w = _door_workload()
    flush("case: door plate, Fibonacci rays 2^18, K = 1")
    ts = TraceScene(w.scene)
    _, st, gk = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                      keys_cap=4096)
    got = set()
    for hi, lo in gk:
        h, b, pl, tile = O.key_fields(hi, lo)
        got.add((h, b, tuple(pl[:bin(b).count("1")]), tile))
    exp, Hx = P.door_expected()
    flush(f"  records gpu {len(got)} by hand {len(exp)}; missing {sorted(exp - got)} extra {sorted(got - exp)}")
    assert got == exp
    res, st2, _ = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize, flags=2)
    torch.cuda.synchronize()
    Hg = to_complex(res)[0]
    err = np.abs(Hg - Hx)
    flush(f"  H per Rx vs closed form: max rel {np.max(err[Hx != 0] / np.abs(Hx[Hx != 0])):.2e}; "
          f"Rx with no path: |H| = {np.abs(Hg[Hx == 0]).max():.1e}")
    assert np.all(err <= 1e-5 * np.abs(Hx))
    sc = O.OracleScene.from_scene(w.scene)
    check_discovery_fib(sc, w, gk)


def test_door_plate_reference_rays(gpu):
    """The reference's TEA/LCG launch (gated, reserve 100) through the door
    plate: the GPU record set is the oracle's and a subset of the
    hand-enumerated one, with transmitted records among them."""
    from automs_amd.tracer import TraceScene, trace
    from oracle import oracle as O
    P, w = _door_workload()
    launch = (256, 256, 2, 100)
    flush(f"case: door plate, reference launch {launch}")
    ts = TraceScene(w.scene)
    _, st, gk = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, 0, False, w.gridsize, keys_cap=4096,
                      reference_launch=launch)
    sc = O.OracleScene.from_scene(w.scene)
    check_discovery_ref(sc, w, launch, gk)
    got = set()
    for hi, lo in gk:
        h, b, pl, tile = O.key_fields(hi, lo)
        got.add((h, b, tuple(pl[:bin(b).count("1")]), tile))
    exp, _ = P.door_expected()
    assert got <= exp
    assert sum(1 for k in got if k[:2] == (1, 0)) >= 4


@pytest.fixture(scope="module")
def door_4321(tmp_path_factory):
    """The 4321 scene with one door leaf (hinge at (-4, 3), 1 m wide, from
    the floor to 2 m up) in the AP's view, opened by 0.3 rad."""
    from automs_amd import channel, geometry
    from automs_amd.scenes import ArraySpec, Workload
    d = tmp_path_factory.mktemp("door4321")
    (d / "door.txt").write_text("-4.0 3.0 0.6\n-4.0 3.0 -1.4\n-4.0 4.0 -1.4\n0\n")
    doors = channel.get_doors(str(d), 0.3)
    sc = channel.load_scene_folder("tests/golden/scene_4321", door_buffer=doors)
    src = geometry.gen_pos_matrix([-3, 3.4, -0.51], [-1, 0, 0], [0, 0, -1], (6, 6), WL / 2).numpy()
    dst = geometry.gen_pos_from_endpoint([-6, 2.4, 42], [-1, 6.5, 40], -0.51).numpy()
    w = Workload("4321+door", sc, ArraySpec(src, np.array([-1, 0, 0], np.float32), 2, 0.5e-5),
                 ArraySpec(dst, np.zeros(3, np.float32), 0, 0.06), WL, 4, 1 << 13, True)
    return w


def _has_transmission(keys):
    from oracle import oracle as O
    n = 0
    for hi, lo in keys:
        h, b, _, _ = O.key_fields(hi, lo)
        n += bin(b).count("1") < h
    return n


def test_door_4321_fibonacci(gpu, door_4321):
    import torch

    from automs_amd.tracer import TraceScene, trace
    from oracle import oracle as O
    w = door_4321
    flush(f"case: 4321 + door, {w.scene.n_tri} triangles, Fibonacci 2^13 per element, K = 4")
    assert w.scene.mat_prop[-1].tolist() == [np.float32(-0.7)] * 19
    ts = TraceScene(w.scene)
    sc = O.OracleScene.from_scene(w.scene)
    _, st, gk = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                      keys_cap=1 << 20)
    ok, ost = check_discovery_fib(sc, w, gk)
    nt = _has_transmission(gk)
    flush(f"  records with a transmission: {nt}")
    assert nt > 10
    Ho, oe, mask = oracle_trace(w, sc, seed=ok)
    res, ge, _ = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                       seed_keys=ok)
    torch.cuda.synchronize()
    r = check_h(to_complex(res), Ho, mask, ge, oe, "4321+door fib:")
    assert r["rel"] < 1e-4


def test_door_4321_reference_rays(gpu, door_4321):
    import torch

    from automs_amd.tracer import TraceScene, trace
    from oracle import oracle as O
    w = door_4321
    launch = (100, 100, 1, 100)
    flush(f"case: 4321 + door, reference launch {launch}, K = 4")
    ts = TraceScene(w.scene)
    sc = O.OracleScene.from_scene(w.scene)
    res, st, gk = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, 0, False, w.gridsize, keys_cap=1 << 20,
                        reference_launch=launch)
    torch.cuda.synchronize()
    ok, ost = check_discovery_ref(sc, w, launch, gk)
    assert _has_transmission(gk) > 0
    Ho, oe, mask = oracle_trace(w, sc, seed=ok)
    if keyset(ok) == keyset(gk):
        r = check_h(to_complex(res), Ho, mask, st, oe, "4321+door ref:")
        assert r["rel"] < 1e-4
