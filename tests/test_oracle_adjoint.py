"""Pins for the oracle's adjoint of the taped MALA recursion (CPU only):
central finite differences of a trajectory statistic through the
decision-frozen recursion, and through the hard-decision sequential MALA of
src/MALA/MALA.jl:316-346 where no decision flips inside the difference step."""
import math

import numpy as np
import pytest

from oracle import deer_oracle as O


def _stat(S):
    # the C4 objective: trajectory mean of |x_t|^2
    return float(np.mean(np.sum(S * S, axis=0)))


def _stat_grad(S):
    return 2.0 * S / S.shape[1]


def _logreg(rng, N, D, s):
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    return Xd, y, O.LogReg(Xd, y, s)


def test_frozen_recursion_equals_sequential_with_its_decisions():
    rng = np.random.default_rng(1)
    Xd, y, tgt = _logreg(rng, 300, 5, 1.3)
    x0 = 0.3 * rng.standard_normal(5)
    Xi, U = O.make_tape(rng, 5, 60)
    S, acc, _ = O.run_mala_sequential_taped(tgt, x0, 0.02, Xi, U)
    assert 0 < acc.sum() < 60
    assert np.array_equal(O.run_mala_frozen(tgt, x0, 0.02, Xi, acc), S)


@pytest.mark.parametrize("scale", [0.5, 2.0])
def test_adjoint_matches_frozen_finite_differences_logreg(scale):
    rng = np.random.default_rng(2)
    D, T, eps = 6, 80, 0.03
    Xd, y, tgt = _logreg(rng, 400, D, scale)
    x0 = 0.2 * rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    S, acc, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    assert 0 < acc.sum() < T
    deps, ds, dx0, _ = O.mala_adjoint(tgt, eps, Xi, x0, S, acc, _stat_grad(S))
    L = lambda e, s, x: _stat(O.run_mala_frozen(O.LogReg(Xd, y, s), x, e, Xi, acc))
    h = 1e-5
    fd_e = (L(eps * (1 + h), scale, x0) - L(eps * (1 - h), scale, x0)) / (2 * h * eps)
    fd_s = (L(eps, scale * (1 + h), x0) - L(eps, scale * (1 - h), x0)) / (2 * h * scale)
    assert abs(deps - fd_e) <= 1e-6 * max(1.0, abs(fd_e))
    assert abs(ds - fd_s) <= 1e-6 * max(1.0, abs(fd_s))
    for d in (0, D - 1):
        e = np.zeros(D)
        e[d] = 1e-5
        fd_x = (L(eps, scale, x0 + e) - L(eps, scale, x0 - e)) / 2e-5
        assert abs(dx0[d] - fd_x) <= 1e-6 * max(1.0, abs(fd_x))


def test_adjoint_matches_hard_decision_differences_when_no_flip():
    """Sequential MALA with hard decisions: where no decision flips within
    +-h, its central difference equals the frozen adjoint."""
    rng = np.random.default_rng(3)
    D, T, eps = 4, 50, 0.05
    tgt = O.DiagGaussian(rng.uniform(0.5, 2.0, D), rng.standard_normal(D))
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    S, acc, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    deps, ds, dx0, _ = O.mala_adjoint(tgt, eps, Xi, x0, S, acc, _stat_grad(S))
    assert ds == 0.0
    h = 1e-7 * eps
    Sp, accp, _ = O.run_mala_sequential_taped(tgt, x0, eps + h, Xi, U)
    Sm, accm, _ = O.run_mala_sequential_taped(tgt, x0, eps - h, Xi, U)
    assert np.array_equal(accp, acc) and np.array_equal(accm, acc)
    fd = (_stat(Sp) - _stat(Sm)) / (2 * h)
    assert abs(deps - fd) <= 1e-5 * max(1.0, abs(fd))


def test_adjoint_dense_gaussian_and_dropped_term():
    """Dense Gaussian: frozen FD agrees; dropping the eps H term of J^T
    (a plausible transposition/omission slip) breaks the agreement."""
    rng = np.random.default_rng(4)
    D, T = 5, 40
    Lam, ev = O.make_corr_gaussian(rng, D, 1.5)
    tgt = O.Gaussian(Lam, rng.standard_normal(D))
    eps = 0.5 / ev.max()
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    S, acc, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    deps, _, dx0, lam = O.mala_adjoint(tgt, eps, Xi, x0, S, acc, _stat_grad(S))
    L = lambda e: _stat(O.run_mala_frozen(tgt, x0, e, Xi, acc))
    h = 1e-6 * eps
    fd = (L(eps + h) - L(eps - h)) / (2 * h)
    assert abs(deps - fd) <= 1e-6 * max(1.0, abs(fd))
    # broken variant: lambda_t = G_t + lambda_{t+1} (no Hessian term)
    G = _stat_grad(S)
    lam_bad = np.cumsum(G[:, ::-1], axis=1)[:, ::-1]
    a = acc.astype(float)[None, :]
    Xbar = O._xbar(x0, S)
    deps_bad = float(np.sum(a * lam_bad * (tgt.grad_batch(Xbar) + Xi / math.sqrt(2 * eps))))
    assert abs(deps_bad - fd) > 1e-3 * abs(fd)
