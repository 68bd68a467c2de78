"""ams_cgemm_frag_noalloc (the contraction without a workspace argument, for
callers that cannot allocate) through the C-ABI against ams_cgemm_frag with
a caller workspace: bit-identical results on two calls in a row, the second
with a smaller input (the scratch keeps the first call's partials); no device
allocation by either the library or torch during the calls; a size beyond
the scratch is refused."""
import ctypes as C

import pytest
import torch

pytestmark = pytest.mark.gpu


def _crand(*shape, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.complex(torch.randn(*shape, generator=g), torch.randn(*shape, generator=g))


def test_noalloc_equals_workspace_variant(gpu):
    from automs_amd import _lib
    from automs_amd.phasemap import frag_pack, split_count
    L = _lib.lib()
    cap = int(L.ams_cgemm_noalloc_max_bytes())
    assert cap == 32 << 20
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for (Ra, Rb, Kd, conj) in ((30, 4096, 1040, False), (16, 1000, 300, True)):
        A = _crand(Ra, Kd, seed=Ra).to(gpu)
        B = _crand(Rb, Kd, seed=Rb).to(gpu)
        fa, fb = frag_pack(A, Ra, Kd, Kd, 1), frag_pack(B, Rb, Kd, Kd, 1)
        splits = max(2, split_count(Ra, Rb, Kd))
        need = int(L.ams_cgemm_workspace_bytes(Ra, Rb, splits))
        assert 0 < need <= cap
        ws = torch.empty(need, dtype=torch.uint8, device=gpu)
        ref = torch.empty(Ra, Rb, dtype=torch.complex64, device=gpu)
        out = torch.full((Ra, Rb), float("nan"), dtype=torch.complex64, device=gpu)
        _lib.check(L.ams_cgemm_frag(C.c_void_p(fa.data_ptr()), C.c_void_p(fb.data_ptr()), Ra, Rb, Kd, int(conj),
                                    C.c_void_p(ref.data_ptr()), Rb, splits, C.c_void_p(ws.data_ptr()), st),
                   "ams_cgemm_frag")
        torch.cuda.synchronize()
        before_lib, before_torch = int(L.ams_mem_outstanding()), torch.cuda.memory_allocated()
        _lib.check(L.ams_cgemm_frag_noalloc(C.c_void_p(fa.data_ptr()), C.c_void_p(fb.data_ptr()), Ra, Rb, Kd,
                                            int(conj), C.c_void_p(out.data_ptr()), Rb, splits, st),
                   "ams_cgemm_frag_noalloc")
        torch.cuda.synchronize()
        assert int(L.ams_mem_outstanding()) == before_lib and torch.cuda.memory_allocated() == before_torch
        print(f"  Ra={Ra} Rb={Rb} Kd={Kd} conj={conj} splits={splits}: {need} B of partials; "
              f"equal: {torch.equal(out, ref)}", flush=True)
        assert torch.equal(out, ref)
    # beyond the scratch: refused with a message, the output left untouched
    Ra, Rb, Kd = 256, 65536, 64
    assert int(L.ams_cgemm_workspace_bytes(Ra, Rb, 2)) > cap
    A = torch.zeros(Ra, Kd, dtype=torch.complex64, device=gpu)
    B = torch.zeros(Rb, Kd, dtype=torch.complex64, device=gpu)
    fa, fb = frag_pack(A, Ra, Kd, Kd, 1), frag_pack(B, Rb, Kd, Kd, 1)
    out = torch.ones(Ra, Rb, dtype=torch.complex64, device=gpu)
    rc = L.ams_cgemm_frag_noalloc(C.c_void_p(fa.data_ptr()), C.c_void_p(fb.data_ptr()), Ra, Rb, Kd, 0,
                                  C.c_void_p(out.data_ptr()), Rb, 2, st)
    assert rc == _lib.AMS_ERR_ARG and b"scratch" in L.ams_last_error()
    torch.cuda.synchronize()
    assert bool((out == 1).all())
