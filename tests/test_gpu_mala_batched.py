"""GPU sequential MALA for many chains (pmcmc_mala_run_batched;
mala_step_batched!, src/MALA/MALA.jl:487-528, test/test-GPU-MALA.jl):

* one step equals the oracle's mala_step_batched on every target family
  (FP64 at 1e-12, identical decisions);
* K steps of N chains equal the oracle's sequential taped runs of each chain
  (FP64 1e-10, identical decisions; FP32 1e-5 relative with the decisions
  away from ties identical).
"""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _model(kind, D, seed):
    from ParallelMCMC import DensityModel, DiagGaussianTarget, GaussianTarget, LogisticTarget

    rng = np.random.default_rng(seed)
    if kind == "diag":
        p = rng.uniform(0.5, 3.0, D)
        m = 0.3 * rng.standard_normal(D)
        return DensityModel(dim=D, target=DiagGaussianTarget(p, m)), O.DiagGaussian(p, m)
    if kind == "gauss":
        Lam, _ = O.make_corr_gaussian(rng, D, 2.0)
        mu = 0.2 * rng.standard_normal(D)
        return DensityModel(dim=D, target=GaussianTarget(Lam, mu)), O.Gaussian(Lam, mu)
    Xd, y, _ = O.make_logreg_data(rng, 400, D)
    return DensityModel(dim=D, target=LogisticTarget(Xd, y, 1.0)), O.LogReg(Xd, y, 1.0)


@pytest.mark.parametrize("kind,D,N", [("diag", 3, 50), ("gauss", 20, 77), ("logistic", 10, 64)])
def test_one_step_matches_oracle(kind, D, N):
    from ParallelMCMC import MALA

    model, tgt = _model(kind, D, 3)
    rng = np.random.default_rng(4)
    eps = 0.05 if kind != "logistic" else 0.01
    X = 0.5 * rng.standard_normal((D, N))
    Xi = rng.standard_normal((D, N))
    U = rng.uniform(0.01, 0.99, N)
    Xn, acc = MALA.mala_step_batched(model, X, eps, Xi, U)
    ref, acc_ref = O.mala_step_batched(tgt, X, eps, Xi, U)
    assert np.array_equal(acc.cpu().numpy().astype(bool), acc_ref)
    assert np.max(np.abs(Xn.cpu().numpy() - ref)) < 1e-12


@pytest.mark.parametrize("kind,D", [("diag", 4), ("gauss", 16), ("logistic", 12)])
def test_many_chains_many_steps_f64(kind, D):
    from ParallelMCMC import MALA

    model, tgt = _model(kind, D, 7)
    rng = np.random.default_rng(8)
    N, K = 40, 200
    eps = 0.1 if kind != "logistic" else 0.02
    X0 = rng.standard_normal((D, N))
    Xi = rng.standard_normal((D, N, K))
    U = rng.random((N, K))
    S, acc = MALA.mala_run_batched(model, X0, eps, Xi, U)
    S, acc = S.cpu().numpy(), acc.cpu().numpy().astype(bool)
    for j in range(0, N, 7):
        ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, X0[:, j], eps, Xi[:, j, :], U[j], check_u=False)
        assert np.array_equal(acc[j], acc_ref)
        assert np.max(np.abs(S[:, j, :] - ref)) < 1e-10


def test_many_chains_f32_logistic():
    from ParallelMCMC import MALA

    model, tgt = _model("logistic", 50, 9)
    rng = np.random.default_rng(10)
    N, K, eps = 64, 128, np.float32(0.005)
    X0 = (0.1 * rng.standard_normal((50, N))).astype(np.float32)
    Xi = rng.standard_normal((50, N, K)).astype(np.float32)
    U = rng.random((N, K)).astype(np.float32)
    S, acc = MALA.mala_run_batched(model, X0, eps, Xi, U)
    S = S.cpu().numpy().astype(np.float64)
    for j in range(0, N, 9):
        ref, _, _ = O.run_mala_sequential_taped(tgt, X0[:, j].astype(np.float64), float(eps),
                                                Xi[:, j, :].astype(np.float64), U[j].astype(np.float64))
        assert O.trace_error(S[:, j, :], ref) < 1e-5
