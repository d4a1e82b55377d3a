"""Object lifetimes at the C ABI: targets and workspaces hold their own
reference to the context they were created on, so a host language whose
finalizers run in no fixed order (Python's cyclic collector at interpreter
exit, Julia's GC) may destroy the context first."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_context_destroyed_before_its_objects():
    import torch

    from ParallelMCMC import _lib

    lib = _lib.load()
    ctx = C.c_void_p()
    assert lib.pmcmc_ctx_create(torch.cuda.current_device(), None, C.byref(ctx)) == 0
    ws = C.c_void_p()
    assert lib.pmcmc_deer_ws_create(ctx, 8, 64, _lib.PREC_F32, C.byref(ws)) == 0
    tgt = C.c_void_p()
    prec = torch.full((8,), 2.0, dtype=torch.float64, device="cuda")
    assert lib.pmcmc_target_diag_gaussian(ctx, 8, _lib.PREC_F64, C.c_void_p(prec.data_ptr()), None,
                                          C.byref(tgt)) == 0
    torch.cuda.synchronize()
    assert lib.pmcmc_ctx_destroy(ctx) == 0
    assert lib.pmcmc_deer_ws_destroy(ws) == 0  # synchronizes the (still live) stream
    assert lib.pmcmc_target_destroy(tgt) == 0  # the last reference: frees the context
    torch.cuda.synchronize()
