"""Pins for the oracle's MALA, target and Jacobian functions (CPU only).

Each test checks the oracle against something it does not itself compute:
closed-form densities, forced decisions, finite differences, brute-force
Jacobians or statistics.  Mirrors test/test-MALA-Kernel.jl,
test/test-Jacobian-Estimator.jl and test/test-Deer-vs-MALA.jl of the reference.
"""
import math

import numpy as np
import pytest
from scipy.stats import multivariate_normal

from oracle import deer_oracle as O


def std_normal(D):
    return O.DiagGaussian.std_normal(D)


def test_logq_matches_gaussian_density():
    # logq is the log density of N(x + eps*g, 2 eps I)  (MALA.jl:172-193)
    rng = np.random.default_rng(1)
    for D in (1, 3, 6):
        x, y, g = rng.standard_normal((3, D))
        eps = 0.4
        ref = multivariate_normal(mean=x + eps * g, cov=2 * eps * np.eye(D)).logpdf(y)
        assert abs(O.logq_mala(y, x, g, eps) - ref) < 1e-12 * max(1, abs(ref))
        # batched form equals the scalar form column by column
        X, Y, G = (rng.standard_normal((D, 5)) for _ in range(3))
        b = O.logq_mala_batched(Y, X, G, eps)
        for t in range(5):
            assert abs(b[t] - O.logq_mala(Y[:, t], X[:, t], G[:, t], eps)) < 1e-12


def test_logalpha_independent_formula():
    # test/test-MALA-Kernel.jl:169-197
    rng = np.random.default_rng(20251231)
    D, eps = 6, 0.40
    tgt = std_normal(D)
    x, xi = rng.standard_normal((2, D))
    y, _ = O.mala_proposal(tgt, x, eps, xi)
    la = O.mala_logalpha(tgt, x, y, eps)
    # independent: densities of N(0,I) and of the two proposal Gaussians
    lp = lambda z: multivariate_normal(np.zeros(D), np.eye(D)).logpdf(z)
    lqxy = multivariate_normal(y - eps * y, 2 * eps * np.eye(D)).logpdf(x)
    lqyx = multivariate_normal(x - eps * x, 2 * eps * np.eye(D)).logpdf(y)
    ref = (lp(y) + lqxy) - (lp(x) + lqyx)
    assert abs(la - ref) < 1e-10


def test_forced_accept_and_reject():
    # test/test-MALA-Kernel.jl:115-167
    rng = np.random.default_rng(20251231)
    D, eps = 5, 0.45
    tgt = std_normal(D)
    x = rng.standard_normal(D)
    for _ in range(2000):
        xi = rng.standard_normal(D)
        y, _ = O.mala_proposal(tgt, x, eps, xi)
        la = O.mala_logalpha(tgt, x, y, eps)
        if la < -1e-8:
            break
    tau = math.exp(la)
    xa, acc, _ = O.mala_step_taped(tgt, x, eps, xi, tau / 2)
    assert acc and np.array_equal(xa, y)
    xr, acc, _ = O.mala_step_taped(tgt, x, eps, xi, (tau + 1) / 2)
    assert (not acc) and np.array_equal(xr, x)
    with pytest.raises(ValueError):
        O.mala_step_taped(tgt, x, eps, xi, 0.0)
    with pytest.raises(ValueError):
        O.mala_step_taped(tgt, x, eps, xi[:-1], 0.5)


def test_sequential_determinism_and_rejections():
    rng = np.random.default_rng(12345)
    D, T, eps = 3, 64, 0.6
    tgt = std_normal(D)
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    S1, a1, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    S2, a2, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    assert np.array_equal(S1, S2) and np.array_equal(a1, a2)
    assert (~a1).sum() > 0  # rejections exercised (test-Deer-vs-MALA.jl:50-69)
    prev = np.concatenate([x0[:, None], S1[:, :-1]], axis=1)
    assert np.array_equal(S1[:, ~a1], prev[:, ~a1])  # rejected => unchanged


def test_stationary_moments_std_normal():
    # test/test-MALA-Kernel.jl:82-113
    rng = np.random.default_rng(20251231)
    D, eps, T, burn = 3, 0.30, 25000, 5000
    tgt = std_normal(D)
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    S, _, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    X = S[:, burn:]
    mu = X.mean(axis=1)
    C = np.cov(X)
    assert np.max(np.abs(mu)) < 0.08
    assert np.max(np.abs(np.diag(C) - 1)) < 0.12
    assert np.max(np.abs(C - np.diag(np.diag(C)))) < 0.10


def _fd_grad(f, x, h=1e-5):
    g = np.zeros_like(x)
    for i in range(len(x)):
        e = np.zeros_like(x)
        e[i] = h
        g[i] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def _targets(rng):
    D = 5
    Lam, _ = O.make_corr_gaussian(rng, D, 2.0)
    Xd, y, _ = O.make_logreg_data(rng, 40, D)
    return [
        std_normal(D),
        O.DiagGaussian(rng.uniform(0.5, 3, D), rng.standard_normal(D)),
        O.Gaussian(Lam, rng.standard_normal(D)),
        O.LogReg(Xd, y, 1.0),
        O.LogReg(Xd, y, 0.7),
    ]


def test_target_grad_hvp_finite_differences():
    rng = np.random.default_rng(3)
    for tgt in _targets(rng):
        x, v = rng.standard_normal((2, tgt.dim))
        g = tgt.grad(x)
        assert np.allclose(g, _fd_grad(tgt.logp, x), rtol=1e-7, atol=1e-7)
        hv_fd = (tgt.grad(x + 1e-5 * v) - tgt.grad(x - 1e-5 * v)) / 2e-5
        assert np.allclose(tgt.hvp(x, v), hv_fd, rtol=1e-6, atol=1e-7)
        # Hessian diagonal vs e_i^T H e_i
        hd = tgt.hess_diag_batch(x[:, None])[:, 0]
        for i in range(tgt.dim):
            e = np.zeros(tgt.dim)
            e[i] = 1
            assert abs(hd[i] - tgt.hvp(x, e)[i]) < 1e-10


def test_logreg_matches_reference_formula():
    # bayes_logreg.jl:37-46 (non-stable but equal form) on moderate logits
    rng = np.random.default_rng(4)
    Xd, y, _ = O.make_logreg_data(rng, 30, 4)
    tgt = O.LogReg(Xd, y)
    b = rng.standard_normal(4)
    lg = Xd @ b
    ref = sum(y[n] * (-math.log1p(math.exp(-lg[n]))) + (1 - y[n]) * (-math.log1p(math.exp(lg[n])))
              for n in range(30)) - 0.5 * float(b @ b)
    assert abs(tgt.logp(b) - ref) < 1e-12
    p = 1 / (1 + np.exp(-lg))
    assert np.allclose(tgt.grad(b), Xd.T @ (y - p) - b, atol=1e-13)


def _surrogate_F(tgt, x0, eps, xi, u):
    """F(x) = a y(x) + (1-a) x + (y0 - x0) sigmoid(logalpha(x) - log u) with a
    and (y0 - x0) frozen at x0: its derivative is the reference's JVP."""
    y0, _ = O.mala_proposal(tgt, x0, eps, xi)
    la0 = O.mala_logalpha(tgt, x0, y0, eps)
    a = 1.0 if math.log(u) < la0 else 0.0

    def F(x):
        y, _ = O.mala_proposal(tgt, x, eps, xi)
        la = O.mala_logalpha(tgt, x, y, eps)
        g = 1 / (1 + math.exp(-(la - math.log(u))))
        return a * y + (1 - a) * x + (y0 - x0) * g

    return F


def test_jvp_matches_central_differences():
    rng = np.random.default_rng(5)
    for tgt in _targets(rng):
        D = tgt.dim
        eps = 0.05 if not isinstance(tgt, O.Gaussian) else 0.01
        for _ in range(3):
            x, xi, v = rng.standard_normal((3, D))
            u = float(rng.uniform(0.05, 0.95))
            jv = O.mala_step_surrogate_sigmoid_jvp(tgt, x, eps, xi, u, v)
            F = _surrogate_F(tgt, x, eps, xi, u)
            h = 1e-6
            fd = (F(x + h * v) - F(x - h * v)) / (2 * h)
            assert np.allclose(jv, fd, rtol=1e-5, atol=1e-6), (type(tgt).__name__, jv, fd)


def test_batched_fwd_jvp_equals_scalar_columns():
    rng = np.random.default_rng(6)
    for tgt in _targets(rng):
        D, T, eps = tgt.dim, 9, 0.1
        X, Xi, Z = rng.standard_normal((3, D, T))
        U = rng.uniform(0.01, 0.99, T)
        fj = O.mala_step_batched_fwd_and_jvp(tgt, X, eps, Xi, U, Z)
        for t in range(T):
            xn, jv = O.mala_step_taped_and_jvp(tgt, X[:, t], eps, Xi[:, t], U[t], Z[:, t])
            assert np.allclose(fj.FT[:, t], xn, atol=1e-12)
            assert np.allclose(fj.Jt[:, t], jv, rtol=1e-9, atol=1e-10)
        Xn, acc = O.mala_step_batched(tgt, X, eps, Xi, U)
        assert np.array_equal(acc, fj.acc) and np.allclose(Xn, fj.FT)


def test_exact_diag_equals_basis_jvps():
    rng = np.random.default_rng(7)
    for tgt in _targets(rng):
        D, T, eps = tgt.dim, 6, 0.08
        X, Xi = rng.standard_normal((2, D, T))
        U = rng.uniform(0.01, 0.99, T)
        FT, Jd, acc, _ = O.mala_jac_diag_batched(tgt, X, eps, Xi, U)
        for t in range(T):
            jvp = lambda x, tt, v: O.mala_step_surrogate_sigmoid_jvp(tgt, x, eps, Xi[:, tt], U[tt], v)
            d = O.jac_diag_via_jvps(jvp, X[:, t], t)
            assert np.allclose(Jd[:, t], d, rtol=1e-10, atol=1e-10)


def test_jac_diag_stoch_mse_decreases():
    # test/test-Jacobian-Estimator.jl:74-106
    rng = np.random.default_rng(20251231)
    D = 8
    A = 0.6 * np.eye(D) + 0.05 * rng.standard_normal((D, D))
    jvp = lambda x, t, v: A @ v
    x = rng.standard_normal(D)

    def mse(K):
        return np.mean([np.mean((O.jac_diag_stoch(jvp, x, 0, K, np.random.default_rng(10000 + 97 * r))
                                 - np.diag(A)) ** 2) for r in range(25)])

    m1, m8, m64 = mse(1), mse(8), mse(64)
    assert m8 < m1 and m64 < m8
    with pytest.raises(ValueError):
        O.jac_diag_stoch(jvp, x, 0, 0, rng)
    # exact on an affine step (test-Jacobian-Estimator.jl:54-72)
    assert np.allclose(O.jac_diag_via_jvps(jvp, x, 0), np.diag(A), atol=1e-12)


def test_rademacher_properties():
    Z = O.rademacher(1234, 3, 0, 100, 10000)
    assert set(np.unique(Z)) == {-1.0, 1.0}
    assert abs(Z.mean()) < 5e-3
    Z2 = O.rademacher(1234, 4, 0, 100, 10000)
    Z3 = O.rademacher(1234, 3, 1, 100, 10000)
    assert abs((Z * Z2).mean()) < 5e-3 and abs((Z * Z3).mean()) < 5e-3
    # neighbouring entries decorrelated
    flat = Z.T.ravel()
    assert abs((flat[1:] * flat[:-1]).mean()) < 5e-3
    # column offset t0 gives the matching slice
    assert np.array_equal(O.rademacher(9, 0, 0, 7, 20)[:, 5:], O.rademacher(9, 0, 0, 7, 15, t0=5))


def test_exact_full_jacobian_equals_basis_jvps():
    rng = np.random.default_rng(8)
    for tgt in _targets(rng):
        D, T, eps = tgt.dim, 4, 0.05
        X, Xi = rng.standard_normal((2, D, T))
        U = rng.uniform(0.01, 0.99, T)
        FT, J, acc, _ = O.mala_jac_full_batched(tgt, X, eps, Xi, U)
        _, Jh, _, _ = O.mala_jac_full_batched(tgt, X, eps, Xi, U, sigmoid=False)
        for t in range(T):
            jvp = lambda x, tt, v: O.mala_step_surrogate_sigmoid_jvp(tgt, x, eps, Xi[:, tt], U[tt], v)
            Jb = O.jac_full_via_jvps(jvp, X[:, t], t)
            assert np.allclose(J[t], Jb, rtol=1e-10, atol=1e-10)
            # decision frozen: the JVP with g(1-g) = 0, i.e. central differences
            # of the taped step wherever the decision does not flip
            fd = np.stack([(O.mala_step_taped(tgt, X[:, t] + 1e-6 * e, eps, Xi[:, t], U[t])[0]
                            - O.mala_step_taped(tgt, X[:, t] - 1e-6 * e, eps, Xi[:, t], U[t])[0]) / 2e-6
                           for e in np.eye(D)], axis=1)
            assert np.allclose(Jh[t], fd, rtol=1e-6, atol=1e-6)
            assert np.allclose(np.diag(J[t]), O.mala_jac_diag_batched(tgt, X, eps, Xi, U)[1][:, t], atol=1e-12)


def test_gaussian_structured_jacobian_form():
    """For H = -Lambda the dense Jacobian (pinned above against basis JVPs)
    is I - a eps Lambda + u c^T with u = g(1-g)(y-x) and
    c = g_y - eps Lambda g_y - g_x - Lambda r + (eps/2) Lambda^2 r,
    r = x - y - eps g_y: the expanded form the GPU's full-Jacobian scan uses."""
    rng = np.random.default_rng(81)
    D, T = 6, 16
    Lam, ev = O.make_corr_gaussian(rng, D, 2.0)
    mu = rng.standard_normal(D)
    tgt = O.Gaussian(Lam, mu)
    eps = 2.0 / ev.max()
    X = mu[:, None] + rng.standard_normal((D, T))
    Xi = rng.standard_normal((D, T))
    U = rng.uniform(0.01, 0.99, T)
    FT, J, acc, la = O.mala_jac_full_batched(tgt, X, eps, Xi, U)
    assert 0 < acc.sum() < T
    for t in range(T):
        x = X[:, t]
        gx = -Lam @ (x - mu)
        y = x + eps * gx + np.sqrt(2 * eps) * Xi[:, t]
        gy = -Lam @ (y - mu)
        r = x - y - eps * gy
        g = 1.0 / (1.0 + np.exp(-(la[t] - np.log(U[t]))))
        c = gy - eps * Lam @ gy - gx - Lam @ r + 0.5 * eps * Lam @ (Lam @ r)
        u = g * (1 - g) * (y - x)
        Js = np.eye(D) - (1.0 if acc[t] else 0.0) * eps * Lam + np.outer(u, c)
        assert np.allclose(Js, J[t], rtol=1e-11, atol=1e-12)


def test_mala_sampler_run_is_the_taped_chain_of_its_own_draws():
    """The MALASampler restatement (src/interface.jl:163-194) draws xi then u
    per step; fed the same draws as a tape, sequential taped MALA gives the
    same states and decisions, and the cached logp equals logp of the state."""
    import numpy as np

    from oracle import deer_oracle as O

    rng = np.random.default_rng(5)
    Xd, y, _ = O.make_logreg_data(rng, 300, 6)
    tgt = O.LogReg(Xd, y)
    x0 = 0.1 * rng.standard_normal(6)
    S, acc, lps = O.mala_sampler_run(tgt, x0, 0.05, 200, np.random.default_rng(11))
    r = np.random.default_rng(11)
    Xi = np.empty((6, 200))
    U = np.empty(200)
    for t in range(200):
        Xi[:, t] = r.standard_normal(6)
        U[t] = r.random()
    S_ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, 0.05, Xi, U)
    assert np.array_equal(acc, acc_ref) and 0 < acc.sum() < 200
    assert np.allclose(S, S_ref, rtol=0, atol=1e-14)
    assert np.allclose(lps, tgt.logp_batch(S), rtol=1e-12, atol=1e-12)
