"""GPU tests of mass-matrix preconditioning (ParallelMALASampler cholM;
src/MALA/MALA.jl:29-47, 430-438, 467-471, 609-614): each block is solved
for theta = L^-1 x on the reparametrized target (standard MALA there with
the same tape is preconditioned MALA on x) and mapped back by x = L theta.

Every block must equal the oracle's sequential preconditioned MALA in x
(run_mala_sequential_taped_mass, restating the reference's cholM branches)
on that block's tape: FP64 within 1e-10 with identical decisions, FP32 within
1e-5 relative; every sample's logp is log p(x) of the model.  The logistic
case runs the dense prior precision L' L / s^2 through the GEMM passes and
the FP64 decisions.
"""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _mass(seed, D):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((D, D))
    M = A @ A.T / D + 0.5 * np.eye(D)
    return np.linalg.cholesky(M)


def _model(kind, D, seed):
    from ParallelMCMC import DensityModel, DiagGaussianTarget, GaussianTarget, LogisticTarget, StdNormalTarget

    rng = np.random.default_rng(seed)
    if kind == "std":
        return DensityModel(dim=D, target=StdNormalTarget(D)), O.DiagGaussian.std_normal(D)
    if kind == "diag":
        p = rng.uniform(0.5, 3.0, D)
        m = 0.3 * rng.standard_normal(D)
        return DensityModel(dim=D, target=DiagGaussianTarget(p, m)), O.DiagGaussian(p, m)
    if kind == "gauss":
        Lam, _ = O.make_corr_gaussian(rng, D, 1.5)
        mu = 0.2 * rng.standard_normal(D)
        return DensityModel(dim=D, target=GaussianTarget(Lam, mu)), O.Gaussian(Lam, mu)
    Xd, y, _ = O.make_logreg_data(rng, 300, D)
    return DensityModel(dim=D, target=LogisticTarget(Xd, y, 0.9)), O.LogReg(Xd, y, 0.9)


def _check_blocks(samples, tgt, sampler, N, L, seed, tol):
    from ParallelMCMC.interface import make_mala_tape_block

    D = L.shape[0]
    got = samples.array()
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(D)
    row = 0
    while row < N:
        Xi, U = make_mala_tape_block(rng, D, sampler.T, sampler.precision)
        rng.integers(0, 2**63 - 1)
        ref, acc, _ = O.run_mala_sequential_taped_mass(tgt, x, sampler.epsilon, Xi.astype(np.float64),
                                                       U.astype(np.float64), L)
        nkeep = min(sampler.T, N - row)
        blk = got[row:row + nkeep].T
        err = np.max(np.abs(blk - ref[:, :nkeep])) / max(1.0, np.max(np.abs(ref)))
        assert err < tol, (row, err)
        x = got[row + nkeep - 1].copy()
        row += nkeep
    lps = np.array([samples[i].logp for i in range(N)])
    return lps, got


@pytest.mark.parametrize("kind,D", [("std", 3), ("diag", 5), ("gauss", 8), ("logistic", 6)])
def test_sampler_mass_matrix_f64(kind, D):
    from ParallelMCMC import ParallelMALASampler, sample

    model, tgt = _model(kind, D, 11 + D)
    L = _mass(5 + D, D)
    eps = 0.02 if kind == "logistic" else 0.15
    N = 150
    sampler = ParallelMALASampler(eps, T=64, tol_abs=1e-13, tol_rel=1e-12, cholM=L)
    samples, infos = sample(np.random.default_rng(31), model, sampler, N, return_infos=True)
    assert all(i.converged for i in infos)
    lps, got = _check_blocks(samples, tgt, sampler, N, L, 31, 1e-10)
    assert np.max(np.abs(lps - tgt.logp_batch(got.T))) < 1e-9


@pytest.mark.parametrize("kind,D", [("gauss", 40), ("logistic", 50)])
def test_sampler_mass_matrix_f32(kind, D):
    from ParallelMCMC import ParallelMALASampler, sample

    model, tgt = _model(kind, D, 3)
    L = _mass(9, D)
    eps = np.float32(0.004 if kind == "logistic" else 0.05)
    N = 256
    sampler = ParallelMALASampler(eps, T=128, tol_abs=1e-6, tol_rel=1e-6, maxiter=400, cholM=L)
    samples = sample(np.random.default_rng(8), model, sampler, N)
    _check_blocks(samples, tgt, sampler, N, L, 8, 1e-5)


def test_cholesky_object_with_L_attribute():
    from ParallelMCMC import ParallelMALASampler

    class Chol:
        def __init__(self, L):
            self.L = L

    L = _mass(1, 3)
    s = ParallelMALASampler(0.1, cholM=Chol(L))
    assert np.array_equal(s._L, L)
