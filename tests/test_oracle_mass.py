"""Pins of the oracle's mass-matrix MALA (src/MALA/MALA.jl:29-47, 172-229,
276-346 with cholM; test/test-MALA-Kernel.jl:199-257 checks the same
identities in the reference):

* logq_mala_mass is the Gaussian log-density N(x + eps M g, 2 eps M);
* with L = I the step and the sequential run are the plain ones;
* u near 0 accepts, u near 1 with a bad proposal rejects (unchanged state);
* the proposal's mean and covariance are x + eps M g and 2 eps M.
"""
import math

import numpy as np
from scipy import stats

from oracle import deer_oracle as O


def _mass(rng, D):
    A = rng.standard_normal((D, D))
    M = A @ A.T / D + 0.5 * np.eye(D)
    return M, np.linalg.cholesky(M)


def test_logq_mass_is_gaussian_density():
    rng = np.random.default_rng(1)
    D, eps = 4, 0.07
    M, L = _mass(rng, D)
    x, y, g = rng.standard_normal((3, D))
    ref = stats.multivariate_normal(mean=x + eps * M @ g, cov=2 * eps * M).logpdf(y)
    assert abs(O.logq_mala_mass(y, x, g, eps, L) - ref) < 1e-12 * max(1, abs(ref))


def test_identity_mass_is_plain_mala():
    rng = np.random.default_rng(2)
    D, T, eps = 5, 40, 0.2
    tgt = O.DiagGaussian.std_normal(D)
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    a, acc_a, la_a = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    b, acc_b, la_b = O.run_mala_sequential_taped_mass(tgt, x0, eps, Xi, U, np.eye(D))
    assert np.allclose(a, b, rtol=0, atol=1e-14) and np.array_equal(acc_a, acc_b)
    assert np.allclose(la_a, la_b, atol=1e-12)


def test_forced_accept_reject():
    rng = np.random.default_rng(3)
    D, eps = 3, 0.3
    tgt = O.DiagGaussian.std_normal(D)
    M, L = _mass(rng, D)
    x = rng.standard_normal(D)
    xi = rng.standard_normal(D)
    _, _, la = O.mala_step_taped_mass(tgt, x, eps, xi, 0.5, L)
    tau = math.exp(min(la, 0.0))
    xn, acc, _ = O.mala_step_taped_mass(tgt, x, eps, xi, tau / 2, L)
    assert acc
    if tau < 1:
        xr, rej, _ = O.mala_step_taped_mass(tgt, x, eps, xi, (tau + 1) / 2, L)
        assert not rej and np.array_equal(xr, x)


def test_proposal_moments():
    rng = np.random.default_rng(4)
    D, eps, n = 3, 0.05, 200_000
    M, L = _mass(rng, D)
    tgt = O.DiagGaussian.std_normal(D)
    x = rng.standard_normal(D)
    Xi = rng.standard_normal((D, n))
    # accept every proposal (u -> 0): the next state is the proposal
    Y = np.stack([O.mala_step_taped_mass(tgt, x, eps, Xi[:, k], 1e-300, L)[0] for k in range(0, n, 50)], axis=1)
    m = x + eps * M @ tgt.grad(x)
    C = np.cov(Y)
    assert np.allclose(Y.mean(axis=1), m, atol=0.02)
    assert np.allclose(C, 2 * eps * M, atol=0.01)
