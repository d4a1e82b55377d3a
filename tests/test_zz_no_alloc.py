"""GPU check of the allocation-free diagonal scan (pmcmc_scan_diag_f32_noalloc,
which has no counterpart in the reference): through the C ABI, its output
equals pmcmc_scan_diag_f32's (same kernels and shapes) and the sequential
recurrence of DEERScan.solve_affine_seq! (src/DEER/DEERScan.jl:51-68, via the
oracle) on two calls in a row, the second smaller, so the shared scratch
array is reused with the first call's contents still in it; a problem whose
scratch exceeds the array is an argument error, after which the entry point
still works."""
import ctypes as C

import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _operands(rng, D, T):
    A = (0.2 * (rng.random((D, T)) - 0.5)).astype(np.float32)
    B = rng.standard_normal((D, T)).astype(np.float32)
    s0 = rng.standard_normal(D).astype(np.float32)
    return A, B, s0


def _run(name, A, B, s0):
    import torch

    from ParallelMCMC import _device, _lib

    dev = torch.device("cuda", 0)
    D, T = A.shape
    A_ = _device.to_colmajor(A, torch.float32, dev)
    B_ = _device.to_colmajor(B, torch.float32, dev)
    s0_ = _device.to_vec(s0, torch.float32, dev)
    S = _device.colmajor_empty(D, T, torch.float32, dev).fill_(float("nan"))
    ctx = _device.context(0)
    torch.cuda.synchronize()
    rc = getattr(_lib.load(), name)(ctx.h, _device.ptr(A_), _device.ptr(B_), _device.ptr(s0_), _device.ptr(S),
                                    C.c_int64(D), C.c_int64(T))
    return rc, S.cpu().numpy(), ctx


@pytest.mark.parametrize("shapes", [((64, 1 << 18), (8, 1000)), ((100, 40000), (3, 17))])
def test_noalloc_scan_matches_allocating_scan(shapes):
    from ParallelMCMC import _lib

    rng = np.random.default_rng(2024 + shapes[0][1])
    for D, T in shapes:  # the second call is smaller than the first
        print(f"noalloc scan D={D} T={T}", flush=True)
        A, B, s0 = _operands(rng, D, T)
        rc_n, S_n, ctx = _run("pmcmc_scan_diag_f32_noalloc", A, B, s0)
        assert rc_n == _lib.PMCMC_OK, _lib.load().pmcmc_last_error(ctx.h)
        rc_a, S_a, _ = _run("pmcmc_scan_diag_f32", A, B, s0)
        assert rc_a == _lib.PMCMC_OK
        assert np.array_equal(S_n, S_a)
        ref = O.solve_affine_seq(A.astype(np.float64), B.astype(np.float64), s0.astype(np.float64))
        assert np.allclose(S_n, ref, rtol=1e-5, atol=1e-5)


def test_noalloc_scan_rejects_oversized_problem_then_still_works():
    from ParallelMCMC import _lib

    rng = np.random.default_rng(7)
    # D = 1024 tiles of 4 columns: 3 * 1024 * T / 4 * 4 bytes of aggregates > 32 MiB at T = 12000
    A, B, s0 = _operands(rng, 1024, 12000)
    rc, _, ctx = _run("pmcmc_scan_diag_f32_noalloc", A, B, s0)
    assert rc == _lib.PMCMC_ERR_ARGUMENT
    assert b"scratch" in _lib.load().pmcmc_last_error(ctx.h)
    A, B, s0 = _operands(rng, 16, 5000)
    rc, S, _ = _run("pmcmc_scan_diag_f32_noalloc", A, B, s0)
    assert rc == _lib.PMCMC_OK
    ref = O.solve_affine_seq(A.astype(np.float64), B.astype(np.float64), s0.astype(np.float64))
    assert np.allclose(S, ref, rtol=1e-5, atol=1e-5)
