"""GPU tests of the DEERScan validation helpers (src/DEER/DEERScan.jl:
solve_affine_seq! :51-68, affine_scan_residual :203-224, check_affine_scan
:240-258; the reference's test/test-DEERScan.jl calls them on small random
problems):

* solve_affine_seq equals the oracle's sequential recurrence bit for bit in
  FP64 and FP32 (one rounding per product and per sum, no FMA, as numpy);
* affine_scan_residual equals the oracle's residual exactly, is 0 on the
  sequential solution, sees a planted error, reports NaN;
* check_affine_scan: ok with tiny errors against the GPU scan on the
  reference's shape list, not ok when the tolerance is below the scan's
  rounding, and the degenerate atol = rtol = 0 case as the reference.
"""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = ((1, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 17), (8, 31), (8, 64), (64, 4096), (257, 1000))


def _problem(seed, D, T, dtype=np.float64):
    rng = np.random.default_rng(seed)
    A = (0.2 * (rng.random((D, T)) - 0.5) + 0.9 * (rng.random((D, T)) < 0.3)).astype(dtype)
    B = rng.standard_normal((D, T)).astype(dtype)
    s0 = rng.standard_normal(D).astype(dtype)
    return A, B, s0


@pytest.mark.parametrize("D,T", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_solve_affine_seq_bitwise(D, T, dtype):
    from ParallelMCMC import DEERScan

    A, B, s0 = _problem(D * 1000 + T, D, T, dtype)
    S = DEERScan.solve_affine_seq(A, B, s0).cpu().numpy()
    ref = O.solve_affine_seq(A, B, s0)
    assert S.dtype == dtype and ref.dtype == dtype
    assert np.array_equal(S, ref)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_affine_scan_residual(dtype):
    from ParallelMCMC import DEERScan

    A, B, s0 = _problem(5, 7, 300, dtype)
    S = O.solve_affine_seq(A, B, s0)
    assert DEERScan.affine_scan_residual(S, A, B, s0) == 0.0
    S2 = S.copy()
    S2[3, 150] += dtype(0.25)
    r = DEERScan.affine_scan_residual(S2, A, B, s0)
    assert r == O.affine_scan_residual(S2, A, B, s0)
    assert r >= 0.24
    S2[1, 0] = np.nan
    assert np.isnan(DEERScan.affine_scan_residual(S2, A, B, s0))
    assert DEERScan.affine_scan_residual(np.zeros((7, 0), dtype), A[:, :0], B[:, :0], s0) == 0.0


@pytest.mark.parametrize("D,T", SHAPES)
def test_check_affine_scan(D, T):
    from ParallelMCMC import DEERScan

    A, B, s0 = _problem(D + 7 * T, D, T)
    res = DEERScan.check_affine_scan(A, B, s0, atol=1e-10, rtol=1e-10)
    assert res["ok"], res["max_abs_err"]
    assert res["max_abs_err"] < 1e-12
    assert res["residual_seq"] == 0.0 and res["residual_scan"] < 1e-12
    assert np.array_equal(res["S_seq"].cpu().numpy(), O.solve_affine_seq(A, B, s0))


def test_check_affine_scan_tolerances():
    from ParallelMCMC import DEERScan

    A, B, s0 = _problem(3, 64, 4096, np.float32)
    loose = DEERScan.check_affine_scan(A, B, s0, atol=1e-5, rtol=1e-5)
    assert loose["ok"]
    tight = DEERScan.check_affine_scan(A, B, s0, atol=0.0, rtol=1e-12)
    assert not tight["ok"] and tight["max_rel_err"] > 1.0
    # atol = rtol = 0: equal entries pass (0 <= 0), their ratio is NaN
    A1, B1, s01 = np.zeros((2, 3)), np.ones((2, 3)), np.zeros(2)
    deg = DEERScan.check_affine_scan(A1, B1, s01, atol=0.0, rtol=0.0)
    assert deg["ok"] and deg["max_abs_err"] == 0.0 and np.isnan(deg["max_rel_err"])
