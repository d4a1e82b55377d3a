"""Dense-Jacobian affine scan S_t = J_t S_{t-1} + B_t (pmcmc_scan_full_*)
against the sequential recurrence (the dense counterpart of
src/DEER/DEERScan.jl:51-68), on shapes that exercise one tile, ragged last
tiles, padded D and one to three levels of the recursive spine; closed forms
J = I (running sums) and J = 0 (S = B); errors."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _seq(J, B, s0):
    D, T = B.shape
    S = np.empty((D, T))
    s = s0.astype(np.float64)
    for t in range(T):
        s = J[:, :, t] @ s + B[:, t]
        S[:, t] = s
    return S


def _problem(rng, D, T, dt):
    # spectral norms ~0.9: errors stay bounded along the scan
    J = (0.45 / np.sqrt(D)) * rng.standard_normal((D, D, T))
    B = rng.standard_normal((D, T))
    s0 = rng.standard_normal(D)
    return J.astype(dt), B.astype(dt), s0.astype(dt)


@pytest.mark.parametrize("D,T", [(1, 1), (3, 5), (17, 64), (17, 65), (64, 1000), (64, 4097), (5, 20000), (33, 600)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_scan_full_matches_sequential(D, T, dt):
    from ParallelMCMC import DEERScan

    rng = np.random.default_rng(D * 100003 + T)
    J, B, s0 = _problem(rng, D, T, dt)
    S = DEERScan.solve_affine_scan_full(J, B, s0).cpu().numpy().astype(np.float64)
    ref = _seq(J.astype(np.float64), B.astype(np.float64), s0.astype(np.float64))
    err = np.max(np.abs(S - ref)) / np.max(np.abs(ref))
    assert err < (1e-12 if dt == np.float64 else 2e-5), err


def test_scan_full_identity_is_running_sum():
    from ParallelMCMC import DEERScan

    rng = np.random.default_rng(2)
    D, T = 40, 3000
    J = np.repeat(np.eye(D)[:, :, None], T, axis=2)
    B = rng.standard_normal((D, T))
    s0 = rng.standard_normal(D)
    S = DEERScan.solve_affine_scan_full(J, B, s0).cpu().numpy()
    ref = s0[:, None] + np.cumsum(B, axis=1)
    assert np.max(np.abs(S - ref)) < 1e-10 * np.max(np.abs(ref))


def test_scan_full_zero_jacobian_returns_B():
    from ParallelMCMC import DEERScan

    rng = np.random.default_rng(3)
    D, T = 64, 777
    B = rng.standard_normal((D, T))
    S = DEERScan.solve_affine_scan_full(np.zeros((D, D, T)), B, rng.standard_normal(D)).cpu().numpy()
    assert np.array_equal(S, B)


def test_scan_full_errors():
    import torch

    from ParallelMCMC import ArgumentError, DEERScan, DimensionMismatch, _device, _lib

    with pytest.raises(DimensionMismatch):
        DEERScan.solve_affine_scan_full(np.zeros((3, 3, 4)), np.zeros((3, 5)), np.zeros(3))
    with pytest.raises(ArgumentError):
        DEERScan.solve_affine_scan_full(np.zeros((65, 65, 2)), np.zeros((65, 2)), np.zeros(65))
    ctx = _device.context(0)
    lib = _lib.load()
    x = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = C.c_void_p(x.data_ptr())
    assert lib.pmcmc_scan_full_f64(ctx.h, p, p, p, C.c_void_p(x.data_ptr() + 8 * 8), 65, 1) == 6  # D > 64
    assert lib.pmcmc_scan_full_f64(ctx.h, p, p, p, p, 2, 0) == 0  # T == 0: no-op
    assert lib.pmcmc_scan_full_f64(ctx.h, p, p, p, p, 0, 3) == 2  # D < 1
    Jb = torch.zeros(2 * 2 * 3, dtype=torch.float64, device="cuda")
    Bb = torch.zeros(2 * 3, dtype=torch.float64, device="cuda")
    rc = lib.pmcmc_scan_full_f64(ctx.h, C.c_void_p(Jb.data_ptr()), C.c_void_p(Bb.data_ptr()), p,
                                 C.c_void_p(Bb.data_ptr()), 2, 3)
    assert rc == 1  # S aliases B
