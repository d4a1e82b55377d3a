"""Pins for the oracle's scans and DEER solver (CPU only).

Mirrors test/test-DEERScan.jl, test/test-Jacobian-Estimator.jl:161-232 and
test/test-Deer-vs-MALA.jl of the reference.
"""
import numpy as np
import pytest

from oracle import deer_oracle as O

SHAPES = ((1, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 17), (8, 31), (8, 64))


def make_affine_problem(rng, D, T):
    # test/test-DEERScan.jl:25-35
    A = 0.2 * (rng.random((D, T)) - 0.5)
    B = rng.standard_normal((D, T))
    s0 = rng.standard_normal(D)
    return A, B, s0


def brute_force_seq(A, B, s0):
    # independent scalar loop
    D, T = A.shape
    S = np.zeros((D, T))
    for d in range(D):
        s = s0[d]
        for t in range(T):
            s = A[d, t] * s + B[d, t]
            S[d, t] = s
    return S


def test_scan_matches_sequential():
    rng = np.random.default_rng(11)
    for D, T in SHAPES:
        A, B, s0 = make_affine_problem(rng, D, T)
        ref = brute_force_seq(A, B, s0)
        assert np.allclose(O.solve_affine_seq(A, B, s0), ref, rtol=1e-12, atol=1e-12)
        assert np.allclose(O.solve_affine_scan_diag(A, B, s0), ref, rtol=1e-12, atol=1e-12)


def test_scan_residual_known_answers():
    # test/test-DEERScan.jl:57-96
    rng = np.random.default_rng(12)
    A, B, s0 = make_affine_problem(rng, 4, 9)
    S = O.solve_affine_seq(A, B, s0)
    assert O.affine_scan_residual(S, A, B, s0) <= 1e-12
    S1 = S.copy()
    S1[1, 0] += 0.25
    assert abs(O.affine_scan_residual(S1, A, B, s0) - 0.25) < 1e-12
    S2 = S.copy()
    S2[2, 5] -= 0.5
    assert abs(O.affine_scan_residual(S2, A, B, s0) - 0.5) < 1e-12
    chk = O.check_affine_scan(A, B, s0, atol=1e-12, rtol=1e-12)
    assert chk["ok"] and chk["max_abs_err"] <= 1e-12
    A0 = np.zeros((4, 0))
    chk0 = O.check_affine_scan(A0, A0, s0)
    assert chk0["ok"] and chk0["max_abs_err"] == 0.0 and chk0["S_scan"].shape == (4, 0)
    with pytest.raises(ValueError):
        O.affine_scan_residual(S, B[:, :4], B, s0)
    with pytest.raises(ValueError):
        O.check_affine_scan(A, B, rng.standard_normal(5))


def test_scan_edge_cases():
    rng = np.random.default_rng(13)
    A, B, s0 = make_affine_problem(rng, 3, 10)
    assert np.allclose(O.solve_affine_scan_diag(np.zeros_like(A), B, s0), B)  # A=0 => S=B
    S = O.solve_affine_scan_diag(A, np.zeros_like(B), s0)  # B=0 => products
    assert np.allclose(S, np.cumprod(A, axis=1) * s0[:, None])


def test_full_seq_reduces_to_diag():
    rng = np.random.default_rng(14)
    A, B, s0 = make_affine_problem(rng, 4, 7)
    J = np.stack([np.diag(A[:, t]) for t in range(7)])
    assert np.allclose(O.solve_affine_seq_full(J, B, s0), O.solve_affine_seq(A, B, s0), atol=1e-14)


def test_batched_hutchinson_exact_on_diagonal_jacobian():
    # test/test-Jacobian-Estimator.jl:161-185: z*(a*z) = a
    rng = np.random.default_rng(20251231)
    D, T = 7, 13
    At = 0.65 + 0.1 * rng.standard_normal((D, T))
    Bt = 0.2 * rng.standard_normal((D, T))
    s0 = rng.standard_normal(D)
    S_in = rng.standard_normal((D, T))
    fwd = lambda Xbar, Z: (At * Xbar + Bt, At * Z, np.ones(T, bool), np.zeros(T))
    S_out, _, _ = O.deer_update_rec(fwd, s0, S_in, "stoch_diag", 1.0, 1, seed=1)
    assert np.allclose(S_out, O.solve_affine_seq(At, Bt, s0), rtol=1e-12, atol=1e-12)


def test_deer_update_argument_validation():
    rng = np.random.default_rng(2)
    D, T = 3, 4
    fwd = lambda Xbar, Z: (0.8 * Xbar + 1, 0.8 * Z, np.ones(T, bool), np.zeros(T))
    S, s0 = rng.standard_normal((D, T)), rng.standard_normal(D)
    with pytest.raises(ValueError):
        O.deer_update_rec(fwd, s0, S, jacobian="full")
    with pytest.raises(ValueError):
        O.deer_update_rec(fwd, s0, S, damping=0.0)
    with pytest.raises(ValueError):
        O.deer_update_rec(fwd, s0, S, probes=0)
    with pytest.raises(ValueError):
        O.deer_solve_rec(fwd, s0, T, maxiter=0)
    with pytest.raises(ValueError):
        O.deer_solve_rec(fwd, s0, T, tol_abs=-1.0)
    with pytest.raises(ValueError):
        O.deer_solve_rec(fwd, s0, T, tol_rel=-1.0)
    with pytest.raises(ValueError):
        O.deer_solve_rec(fwd, s0, T, init=rng.standard_normal((D, T + 1)))


def test_undamped_deer_exact_within_T_iterations():
    """Undamped DEER fixes one more step per iteration for ANY Jacobian
    estimate (column t of S_new is f_t(s_{t-1}) once column t-1 is exact),
    so it reaches the sequential trajectory in at most T+1 iterations."""
    rng = np.random.default_rng(21)
    D, T, eps = 3, 12, 0.35
    tgt = O.DiagGaussian.std_normal(D)
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    S, info = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-300, tol_rel=0.0, maxiter=T + 2,
                           jacobian="stoch_diag", damping=1.0, seed=5)
    assert info.iters <= T + 2
    assert np.allclose(S, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("jac,probes,eps,seed", [("diag", 1, 0.05, 12345), ("stoch_diag", 2, 0.15, 777)])
def test_deer_recovers_sequential_mala(jac, probes, eps, seed):
    # test/test-Deer-vs-MALA.jl:71-108 and :161-215
    rng = np.random.default_rng(seed)
    D, T = 3, 32
    tgt = O.DiagGaussian.std_normal(D)
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    S, info = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-10, tol_rel=1e-8, maxiter=400,
                           jacobian=jac, damping=0.5, probes=probes, seed=42)
    assert info.converged
    assert np.allclose(S, ref, rtol=1e-6, atol=1e-8)
    assert np.array_equal(info.accepted, acc)


def test_deer_defect_logreg():
    # test/test-Jacobian-Estimator.jl:187-232, on the logistic target
    rng = np.random.default_rng(22)
    Xd, y, _ = O.make_logreg_data(rng, 200, 6)
    tgt = O.LogReg(Xd, y)
    D, T, eps = 6, 80, 0.02
    x0 = rng.standard_normal(D) * 0.3
    Xi, U = O.make_tape(rng, D, T)
    S, info = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-12, tol_rel=1e-10, maxiter=300,
                           jacobian="stoch_diag", damping=0.5, seed=3)
    assert info.converged
    prev = x0
    for t in range(T):
        nxt, _, _ = O.mala_step_taped(tgt, prev, eps, Xi[:, t], U[t])
        assert np.linalg.norm(S[:, t] - nxt) <= 1e-9
        prev = S[:, t]


@pytest.mark.parametrize("jac", ["full", "full_sigmoid"])
def test_full_jacobian_deer_on_correlated_gaussian(jac):
    """With a dense Jacobian, undamped DEER is Newton's method on the taped
    recursion (accept decisions frozen per iteration): it reaches the
    sequential trajectory in a few iterations where the exact diagonal needs
    O(T) (eps * lambda_max = 1, condition number 1e3)."""
    rng = np.random.default_rng(13)
    D, T = 16, 128
    Lam, ev = O.make_corr_gaussian(rng, D, 3.0)
    tgt = O.Gaussian(Lam, np.zeros(D))
    eps = 1.0 / ev.max()
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    S, info = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=100, jacobian=jac,
                           damping=1.0)
    assert info.converged and info.iters < 30, info.iters
    assert np.max(np.abs(S - ref)) < 1e-10
    assert np.array_equal(info.accepted, acc)
    _, info_d = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=400,
                             jacobian="diag", damping=1.0, seed=1)
    assert info_d.iters > 3 * info.iters
