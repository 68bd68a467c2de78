"""Device-memory bookkeeping of libams.so (the counterpart of the reference's
cmmCheckMem leak report after every trace, cudaMemManager.h:75-81, called at
optixPathTracer.cpp:857): ams_mem_outstanding() counts the bytes the library
holds; a scene handle's BVH and scratch are released by ams_scene_destroy,
also after a trace that failed part way."""
import numpy as np
import pytest

from _parity import flush

pytestmark = pytest.mark.gpu


def test_outstanding_returns_to_baseline(gpu):
    import torch

    from automs_amd import _lib, scenes
    from automs_amd.tracer import TraceScene, trace
    L = _lib.lib()
    torch.cuda.synchronize()
    base = int(L.ams_mem_outstanding())
    w = scenes.room_boxes(fib_n=1 << 10, n_rx=(16, 16))
    for rep in range(2):
        ts = TraceScene(w.scene)
        after_create = int(L.ams_mem_outstanding())
        assert after_create > base
        trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize)
        trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, 0, False, w.gridsize, reference_launch=(32, 32, 1, 8))
        torch.cuda.synchronize()
        after_trace = int(L.ams_mem_outstanding())
        assert after_trace > after_create  # the handle keeps its scratch for the next trace
        # a trace that fails after its scratch is allocated: a seed key whose
        # tile is outside the Rx grid is refused on the host
        bad = np.array([[0, 0xFFFFF]], np.uint64)
        with pytest.raises(RuntimeError):
            trace(ts, w.tx, w.rx, w.wave_len, w.max_reflection, w.fib_n, w.per_element, w.gridsize, seed_keys=bad)
        torch.cuda.synchronize()
        ts.close()
        end = int(L.ams_mem_outstanding())
        flush(f"  rep {rep}: baseline {base} B, scene {after_create - base} B, after traces {after_trace - base} B, "
              f"after destroy {end - base} B")
        assert end == base
