"""GPU parity of the channel tracer against the FP64 oracle.

Runs the HIP path through the C-ABI (libams.so) and the oracle on the same
seeded inputs at sizes the oracle finishes in seconds.  The bars are the
tie-attributed ones of tests/_parity.py:
  * discovery: every record of a tie-free ray is found, and the record sets
    differ in at most as many records as the oracle counts tie rays;
  * evaluation from identical task lists: every (Tx, Rx) entry without a
    near-tie chain within 1e-4 of its row's largest term, successes equal up
    to the counted tie chains, H within 1e-4 relative in Frobenius norm;
  * coverage 20 log10 |sum_tx H| within 0.05 dB wherever it is above
    -120 dB, except at Rx that hold a tie entry.
"""
import sys

import numpy as np
import pytest

from _parity import check_discovery_fib, check_discovery_ref, check_h, keyset, oracle_trace, to_complex

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_room():
    from automs_amd import scenes
    return scenes.room_boxes(fib_n=1 << 12)


@pytest.fixture(scope="module")
def oracle_scene(small_room):
    from oracle import oracle as O
    return O.OracleScene.from_scene(small_room.scene)


@pytest.fixture(scope="module")
def trace_scene(small_room, gpu):
    from automs_amd.tracer import TraceScene
    return TraceScene(small_room.scene)


def _keyset(k):
    return set(map(tuple, np.asarray(k, dtype=np.uint64).tolist()))


def test_plane_ids_match(small_room, oracle_scene, trace_scene):
    _flush("case: plane ids room_boxes")
    assert np.array_equal(oracle_scene.plane_ids(), trace_scene.plane_ids())


def test_discovery_records(small_room, oracle_scene, trace_scene):
    from automs_amd.tracer import trace
    w = small_room
    _flush(f"case: discovery room_boxes fib_n={w.fib_n}")
    _, st, gk = trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                      w.gridsize, keys_cap=1 << 20)
    ok, ost = check_discovery_fib(oracle_scene, w, gk)
    assert st["rays"] == ost["rays"]


@pytest.mark.parametrize("flags", [2, 0])
def test_evaluate_from_oracle_records(small_room, oracle_scene, trace_scene, flags):
    """Same (path, tile) seed records on both sides: H and coverage parity."""
    import torch

    from automs_amd.tracer import trace
    from oracle import oracle as O
    w = small_room
    _flush(f"case: seeded evaluation room_boxes flags={flags}")
    keys, _ = O.discover(oracle_scene, w.tx.pos, w.rx.pos, w.rx.radius, w.max_reflection, w.gridsize, w.fib_n,
                         w.per_element)
    Ho, ost, mask = oracle_trace(w, oracle_scene, flags=flags, seed=keys)
    res, st, _ = trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                       flags=flags, seed_keys=keys)
    torch.cuda.synchronize()
    r = check_h(to_complex(res), Ho, mask, st, ost, f"room_boxes seeded flags={flags}:")
    assert r["rel"] < 1e-4
    assert abs(st["tasks"] - ost["tasks"]) <= 4 * ost["bwd_ties"]


def test_end_to_end_coverage(small_room, oracle_scene, trace_scene):
    """Discovery, rounds and evaluation on both sides from the same rays."""
    import torch

    from automs_amd.tracer import trace
    w = small_room
    _flush("case: end-to-end room_boxes")
    res, st, gk = trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                        w.gridsize, keys_cap=1 << 20)
    torch.cuda.synchronize()
    ok, _ = check_discovery_fib(oracle_scene, w, gk)
    Ho, ost, mask = oracle_trace(w, oracle_scene)
    if keyset(ok) == keyset(gk):
        r = check_h(to_complex(res), Ho, mask, st, ost, "room_boxes end to end:")
        assert r["rel"] < 1e-4


def test_torch_op_scene_4321_matches_oracle(gpu):
    """The drop-in op torch.ops.pathtrace.torch_launch_pathtrace on the
    reference's own 4321 scene and go.py's __main__ geometry (6x6 AP at
    (-3, 3.4, -0.51) facing -x, 84 x 75 Rx at z = -0.51, 4 reflections,
    100 x 100 launch with the reference's ray set), against the oracle on the
    same inputs and rays."""
    import torch

    from automs_amd import channel, geometry
    from automs_amd.scenes import ArraySpec, Workload
    from oracle import oracle as O
    _flush("case: torch op scene 4321")
    channel.register_torch_op()
    wl = 3e8 / 60e9
    sc = channel.load_scene_folder("tests/golden/scene_4321")
    src = geometry.gen_pos_matrix([-3, 3.4, -0.51], [-1, 0, 0], [0, 0, -1], (6, 6), wl / 2)
    dst = geometry.gen_pos_from_endpoint([-6, 2.4, 84], [-1, 6.5, 75], -0.51)
    result = torch.zeros(36, 84 * 75, 2, dtype=torch.float32, device=gpu)
    tri = torch.tensor(sc.tri.reshape(-1, 3), device=gpu)
    torch.ops.pathtrace.torch_launch_pathtrace(
        result, src.to(gpu), dst.to(gpu), torch.cat([src - 0.5e-5, src + 0.5e-5], -1).to(gpu),
        torch.cat([dst - 0.06, dst + 0.06], -1).to(gpu), tri, torch.tensor(sc.mat, device=gpu), 1, 100, 100, wl, 4,
        100, torch.tensor(sc.mat_prop, device=gpu), torch.tensor([-1.0, 0, 0]), torch.tensor([0.0, 0, 0]), 0, 2, 0, 5,
        0.5e-5, 0.06)
    torch.cuda.synchronize()
    Hg = result.double().cpu().numpy()
    Hg = Hg[..., 0] + 1j * Hg[..., 1]
    # a second call on the same scene reuses the cached scene handle
    n_cached = len(channel._SCENE_CACHE)
    result2 = torch.zeros_like(result)
    torch.ops.pathtrace.torch_launch_pathtrace(
        result2, src.to(gpu), dst.to(gpu), torch.cat([src - 0.5e-5, src + 0.5e-5], -1).to(gpu),
        torch.cat([dst - 0.06, dst + 0.06], -1).to(gpu), tri, torch.tensor(sc.mat, device=gpu), 1, 100, 100, wl, 4,
        100, torch.tensor(sc.mat_prop, device=gpu), torch.tensor([-1.0, 0, 0]), torch.tensor([0.0, 0, 0]), 0, 2, 0, 5,
        0.5e-5, 0.06)
    torch.cuda.synchronize()
    assert len(channel._SCENE_CACHE) == n_cached
    assert torch.equal(result2, result)
    w = Workload("4321", sc, ArraySpec(src.numpy(), np.array([-1, 0, 0], np.float32), 2, 0.5e-5),
                 ArraySpec(dst.numpy(), np.zeros(3, np.float32), 0, 0.06), wl, 4, 100 * 100, False)
    osc = O.OracleScene.from_scene(sc)
    # the op's records: the same trace through the C-ABI with its key output
    from automs_amd.tracer import TraceScene, trace
    launch = (100, 100, 1, 100)
    res2, st2, gk = trace(TraceScene(sc), w.tx, w.rx, wl, 4, 0, False, 5, keys_cap=1 << 20,
                          reference_launch=launch)
    torch.cuda.synchronize()
    H2 = to_complex(res2)
    assert np.linalg.norm(H2 - Hg) <= 1e-6 * np.linalg.norm(Hg)
    keys, _ = check_discovery_ref(osc, w, launch, gk)
    Ho, ost, mask = oracle_trace(w, osc, seed=keys)
    _flush(f"  oracle paths={ost['paths']} tasks={ost['tasks']} ties={ost['bwd_ties']}")
    assert ost["paths"] > 10
    if keyset(keys) == keyset(gk):
        r = check_h(Hg, Ho, mask, st2, ost, "4321 torch op:")
        assert r["rel"] < 1e-4


def test_reference_discovery_mode(small_room, oracle_scene, trace_scene):
    """The reference's own ray set (TEA/LCG per launch index, per-tile
    min-hit gating, reserve cap): tie-attributed record sets, then the
    end-to-end H from those records with the evaluation bars."""
    import torch

    from automs_amd.tracer import trace
    w = small_room
    launch = (100, 100, 2, 100)
    _flush(f"case: reference-mode discovery room_boxes launch={launch}")
    res, st, gk = trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, 0, False, w.gridsize,
                        keys_cap=1 << 20, reference_launch=launch)
    torch.cuda.synchronize()
    ok, ost = check_discovery_ref(oracle_scene, w, launch, gk)
    assert st["rays"] == ost["rays"] == 100 * 100 * 2
    Ho, oe, mask = oracle_trace(w, oracle_scene, seed=ok)
    if keyset(ok) == keyset(gk):
        r = check_h(to_complex(res), Ho, mask, st, oe, "room_boxes reference mode:")
        assert r["rel"] < 1e-4


def test_reference_mode_rejects_bad_launch(trace_scene, small_room):
    from automs_amd.tracer import trace
    w = small_room
    with pytest.raises(RuntimeError):
        trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, 0, False, w.gridsize,
              reference_launch=(0, 10, 1, 10))


@pytest.mark.parametrize("jitter", [0.0, 0.06])
def test_rx_lattice_query_matches_hierarchy(small_room, trace_scene, jitter):
    """The lattice sphere query (regular Rx arrays) returns the hits of the
    sphere hierarchy (flags bit 5): identical record sets and H, also for Rx
    positions jittered off the lattice by up to 6% of the spacing."""
    import torch

    from automs_amd.scenes import ArraySpec
    from automs_amd.tracer import trace
    w = small_room
    _flush(f"case: rx lattice vs hierarchy jitter={jitter}")
    pos = np.asarray(w.rx.pos, np.float32).copy()
    if jitter:
        sp = float(np.linalg.norm(pos[1, 0] - pos[0, 0]))
        pos += (np.random.default_rng(3).uniform(-1, 1, pos.shape) * jitter * sp / np.sqrt(3)).astype(np.float32)
    rx = ArraySpec(pos, w.rx.dir, w.rx.bptype, w.rx.radius)
    out = {}
    for flags in (0, 32):
        res, st, keys = trace(trace_scene, w.tx, rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                              w.gridsize, flags=flags, keys_cap=1 << 20)
        torch.cuda.synchronize()
        out[flags] = (res.double().cpu().numpy(), _keyset(keys), st)
    (Ha, ka, sa), (Hb, kb, sb) = out[0], out[32]
    rel = np.linalg.norm(Ha - Hb) / np.linalg.norm(Hb)
    _flush(f"  keys lattice={len(ka)} hierarchy={len(kb)} symdiff={len(ka ^ kb)} H rel={rel:.2e} "
           f"sphere tests {sa['fwd_spheres']} vs {sb['fwd_spheres']}, nodes {sa['fwd_nodes']} vs {sb['fwd_nodes']}")
    assert ka == kb
    assert rel < 1e-6
    assert sa["fwd_nodes"] < sb["fwd_nodes"]


@pytest.mark.parametrize("flags", [128, 256, 8])
def test_culling_leaves_results_unchanged(small_room, trace_scene, flags):
    """Early termination of branches that can no longer reach the Rx array
    (disabled by flags bit 7) and the analytic pre-check of backward chains
    (bit 8) only skip queries whose outcome is known: identical record sets,
    success counts and H.  Keeping the scene hit across Rx hits (bit 3
    re-queries) may differ only for Rx within 10 um of a wall: none here."""
    import torch

    from automs_amd.tracer import trace
    w = small_room
    _flush(f"case: culling flags={flags}")
    out = {}
    for f in (0, flags):
        res, st, keys = trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                              w.gridsize, flags=f, keys_cap=1 << 20)
        torch.cuda.synchronize()
        out[f] = (res.double().cpu().numpy(), _keyset(keys), st)
    (Ha, ka, sa), (Hb, kb, sb) = out[0], out[flags]
    rel = np.linalg.norm(Ha - Hb) / np.linalg.norm(Hb)
    _flush(f"  keys {len(ka)} vs {len(kb)} successes {sa['successes']} vs {sb['successes']} queries "
           f"{sa['fwd_queries'] + sa['bwd_queries']} vs {sb['fwd_queries'] + sb['bwd_queries']} H rel {rel:.1e}")
    assert ka == kb
    assert sa["successes"] == sb["successes"]
    assert rel < 1e-6
    assert sa["fwd_queries"] + sa["bwd_queries"] <= sb["fwd_queries"] + sb["bwd_queries"]


@pytest.mark.parametrize("mode", ["fibonacci", "reference", "ms"])
def test_device_rounds_match_host_rounds(small_room, trace_scene, mode):
    """The neighbour-expansion rounds built on the device (flags bit 13: the
    expansion over the traced-pair bitmap, the grouping by tile and the units
    are kernels; the default for large maps, such as the office's) against
    the host loop (flags bit 12): the same rounds, tasks, successes and record
    sets, and H equal to FP64 accumulation order (within a tile the device
    orders tasks by path, the host by discovery)."""
    import torch

    from automs_amd import scenes
    from automs_amd.tracer import TraceScene, trace
    _flush(f"case: device-built rounds vs host loop ({mode})")
    if mode == "ms":
        w = scenes.room_boxes_ms(fib_n=1 << 12, ms_fib_n=1 << 16)["F"]
        ts = TraceScene(w.scene)
    else:
        w, ts = small_room, trace_scene
    kw = dict(reference_launch=(64, 64, 1, 8)) if mode == "reference" else {}
    fib = 0 if mode == "reference" else w.fib_n
    out = {}
    for f in (8192, 4096):
        res, st, keys = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, fib, w.per_element, w.gridsize, flags=f,
                              keys_cap=1 << 20, **kw)
        torch.cuda.synchronize()
        out[f] = (res.double().cpu().numpy(), _keyset(keys), st)
    (Ha, ka, sa), (Hb, kb, sb) = out[8192], out[4096]
    rel = np.linalg.norm(Ha - Hb) / max(np.linalg.norm(Hb), 1e-300)
    _flush(f"  rounds {sa['rounds']} vs {sb['rounds']}, tasks {sa['tasks']} vs {sb['tasks']}, successes "
           f"{sa['successes']} vs {sb['successes']}, H rel {rel:.1e}")
    assert sa["rounds"] == sb["rounds"] and sa["tasks"] == sb["tasks"]
    assert sa["successes"] == sb["successes"] and sa["chains"] == sb["chains"]
    assert ka == kb
    assert rel < 1e-6
    if mode == "ms":
        ts.close()


@pytest.mark.parametrize("base", [2048, 16, 16 | 2048])
def test_four_wide_walk_matches_binary(small_room, trace_scene, base):
    """The four-wide hierarchy walk returns the binary walk's hits (flags bit
    10 forces the binary walk), with the scene staged in LDS (bit 11 forces
    the four-wide walk there) and read from global memory (bit 4, where it is
    the default for forward queries; with bit 11 the backward queries walk it
    too): identical record sets and successes, the same H, and fewer node
    boxes tested forward."""
    import torch

    from automs_amd.tracer import trace
    w = small_room
    _flush(f"case: four-wide vs binary walk base flags={base}")
    out = {}
    binary = (base & ~2048) | 1024
    for f in (base, binary):
        res, st, keys = trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element,
                              w.gridsize, flags=f, keys_cap=1 << 20)
        torch.cuda.synchronize()
        out[f] = (res.double().cpu().numpy(), _keyset(keys), st)
    (Ha, ka, sa), (Hb, kb, sb) = out[base], out[binary]
    rel = np.linalg.norm(Ha - Hb) / np.linalg.norm(Hb)
    _flush(f"  keys {len(ka)} vs {len(kb)} successes {sa['successes']} vs {sb['successes']} nodes "
           f"{sa['fwd_nodes']} vs {sb['fwd_nodes']} H rel {rel:.1e}")
    assert ka == kb
    assert sa["successes"] == sb["successes"]
    assert rel < 1e-6
    assert sa["fwd_nodes"] < sb["fwd_nodes"]


def test_trace_many_matches_sequential(gpu):
    """trace_many runs independent traces from separate host threads on
    separate streams: every counter and H, bit for bit, equal those of
    calling trace() on each job in turn (H is summed in fixed point, so the
    order the units finish in does not change it; three metasurface traces,
    reduced ray counts); jobs sharing a scene handle are refused."""
    import torch

    from automs_amd import scenes
    from automs_amd.tracer import TraceScene, trace, trace_many
    _flush("case: trace_many vs sequential (room_boxes_ms, fib 2^14, ms fib 2^18)")
    tr = scenes.room_boxes_ms(fib_n=1 << 14, ms_fib_n=1 << 18)
    hs = {k: TraceScene(w.scene) for k, w in tr.items()}

    def job(k):
        w = tr[k]
        return dict(ts=hs[k], tx=w.tx, rx=w.rx, wave_len=w.wave_len, max_reflection=w.max_reflection,
                    fib_n=w.fib_n, per_element=w.per_element, gridsize=w.gridsize)
    seq = {k: trace(**job(k)) for k in tr}
    torch.cuda.synchronize()
    for rep in range(2):
        par = dict(zip(tr, trace_many([job(k) for k in tr])))
        torch.cuda.synchronize()
        for k in tr:
            assert torch.equal(seq[k][0], par[k][0]), (rep, k)
            for f in ("records_unique", "paths", "tasks", "chains", "successes", "fwd_queries", "bwd_queries"):
                assert seq[k][1][f] == par[k][1][f], (rep, k, f)
        _flush(f"  rep {rep}: identical counters and H for {list(tr)}")
    lim = dict(zip(tr, trace_many([job(k) for k in tr], max_concurrent=2)))  # three jobs on two workers
    torch.cuda.synchronize()
    for k in tr:
        assert lim[k][1]["successes"] == seq[k][1]["successes"], k
    with pytest.raises(ValueError):
        trace_many([job("G"), job("G")])


def test_task_export_rebuilds_H(small_room, trace_scene, tmp_path):
    """The task export (the reference's path.dat under EXPORT_TASKINFO,
    optixPathTracer.cpp:244-256): one 52-byte record per backward chain of
    every round, [src][task][Rx in tile]; the successful records equal the
    success counter, and H rebuilt from them alone (beam patterns of the two
    cosines x |reflectivity| of every recorded hit x loss x e^{-j phase})
    matches the traced H.  The phase is an unwrapped FP32 value (about 1e4
    rad here), so the rebuild agrees to 3e-3.  Exporting leaves H unchanged."""
    import torch

    from automs_amd.tracer import read_task_info, trace
    from oracle import oracle as O
    w = small_room
    _flush("case: task export rebuilds H (room_boxes, fib 2^12)")
    path = tmp_path / "path.dat"
    res0, st0, _ = trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize)
    res, st, _ = trace(trace_scene, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                       task_info_path=str(path))
    torch.cuda.synchronize()
    H0 = res0.double().cpu().numpy()
    H = res.double().cpu().numpy()
    assert np.linalg.norm(H - H0) <= 1e-6 * np.linalg.norm(H0)
    rec = read_task_info(str(path))
    assert len(rec) == st["tasks"] * w.tx.n * w.gridsize ** 2
    ok = rec[rec["successful"] == 1]
    assert len(ok) == st["successes"]
    assert (ok["hit_cnt"] <= w.max_reflection).all()

    def bp(c, t):
        if t == 0:
            return np.ones(len(c))
        return np.array([O.beam_pattern([float(x), float(np.sqrt(max(0.0, 1.0 - float(x) ** 2))), 0.0],
                                        [1.0, 0.0, 0.0], t) for x in c])
    gam = np.abs(np.asarray(w.scene.mat_prop, np.float64).reshape(-1))
    srcs = [0, w.tx.n - 1]  # rebuild the rows of two Tx elements
    sub = ok[np.isin(ok["src"], srcs)]
    hit = np.arange(5)[None, :] < sub["hit_cnt"][:, None]
    refl = np.where(hit, gam[np.where(hit, sub["hit_loss_idx"], 0)], 1.0).prod(axis=1)
    amp = bp(sub["receive_angle"], w.rx.bptype) * bp(sub["launch_angle"], w.tx.bptype) * sub["loss"] * refl
    Hr = np.zeros(H.shape[:2], np.complex128)
    np.add.at(Hr, (sub["src"], sub["dst"]), amp * np.exp(-1j * sub["phase"].astype(np.float64)))
    Hc = (H[..., 0] + 1j * H[..., 1])[srcs]
    rel = np.linalg.norm(Hr[srcs] - Hc) / np.linalg.norm(Hc)
    _flush(f"  records {len(rec)}, successful {len(ok)}, rebuilt {len(sub)} of Tx {srcs}: H rel {rel:.2e}")
    assert rel < 3e-3


@pytest.mark.parametrize("which", ["room", "office"])
def test_slab_bounded_last_leg(small_room, trace_scene, which):
    """Discovery branches with no reflection left query the scene only where
    the ray is in the Rx lattice's slab (flags bit 18 queries the whole ray):
    outside it the pass-through hits cannot change the records.  The room:
    identical record sets and H.  The office (5 reflections): the record sets
    may differ only where a sphere sits at the acceptance boundary (the
    FP32 ray origin is then the reflection point instead of the last
    pass-through point), at most one record in 10,000; H within 1e-6."""
    import torch

    from automs_amd import scenes
    from automs_amd.tracer import TraceScene, trace
    _flush(f"case: slab-bounded last leg, {which}")
    if which == "room":
        w, ts = small_room, trace_scene
    else:
        w = scenes.office(fib_n=1 << 18, n_rx=(80, 60), ap=(2, 2))["H_d"]
        ts = TraceScene(w.scene)
    out = {}
    for f in (0, 262144):
        res, st, keys = trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                              flags=f, keys_cap=1 << 22)
        torch.cuda.synchronize()
        out[f] = (res.double().cpu().numpy(), _keyset(keys), st)
    (Ha, ka, sa), (Hb, kb, sb) = out[0], out[262144]
    rel = np.linalg.norm(Ha - Hb) / np.linalg.norm(Hb)
    _flush(f"  keys {len(ka)} vs {len(kb)}, symmetric difference {len(ka ^ kb)}; forward queries {sa['fwd_queries']} vs "
           f"{sb['fwd_queries']}, node visits {sa['fwd_nodes']} vs {sb['fwd_nodes']}; H rel {rel:.1e}")
    assert sa["fwd_nodes"] < sb["fwd_nodes"]
    if which == "room":
        assert ka == kb and np.array_equal(Ha, Hb)
    else:
        ts.close()
        assert len(ka ^ kb) <= max(2, len(kb) // 10000)
        assert rel < 1e-6


def test_discovery_in_two_launches(gpu):
    """2^25 Fibonacci rays (16 AP elements x 2^21) are discovered in two
    launches of 2^24 rays whose unique records are merged on the device.
    Records depend on each ray alone, so the set equals the union of the
    sets of the AP's two halves (rows 0-1 and 2-3, 2^24 rays each, one
    launch), and it comes back sorted and free of duplicates."""
    from automs_amd import scenes
    from automs_amd.scenes import ArraySpec
    from automs_amd.tracer import TraceScene, trace
    w = scenes.room_boxes(fib_n=1 << 21)
    _flush(f"case: discovery in two launches, room_boxes, {w.tx.n} x 2^21 rays")
    ts = TraceScene(w.scene)

    def run(tx):
        _, st, keys = trace(ts, tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize,
                            flags=2, keys_cap=1 << 22)
        return st, np.asarray(keys, dtype=np.uint64)

    st, full = run(w.tx)
    halves = [run(ArraySpec(np.ascontiguousarray(w.tx.pos[r0:r0 + 2]), w.tx.dir, w.tx.bptype, w.tx.radius))
              for r0 in (0, 2)]
    ts.close()
    _flush(f"  launches {st['n_discover_launches']}, records {len(full)}; halves "
           f"{[h[0]['n_discover_launches'] for h in halves]} launches, {[len(h[1]) for h in halves]} records")
    assert st["n_discover_launches"] == 2 and all(h[0]["n_discover_launches"] == 1 for h in halves)
    assert len(full) == st["records_unique"]
    pairs = [tuple(k) for k in full.tolist()]
    assert pairs == sorted(set(pairs))
    assert set(pairs) == set(map(tuple, halves[0][1].tolist())) | set(map(tuple, halves[1][1].tolist()))
