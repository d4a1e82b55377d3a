"""GPU parity of the dense-Gaussian (precision-matrix) target: MFMA Lambda
GEMM passes vs the oracle — batched MALA step + JVP, log-densities, DEER
solves (FP64 plain path at the 1e-10 bar; FP32 / BF16-Jacobian MFMA path at
the 1e-5 relative bar)."""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _setup(seed, D, cond=2.0, with_mean=True):
    from ParallelMCMC import DensityModel, GaussianTarget

    rng = np.random.default_rng(seed)
    Lam, _ = O.make_corr_gaussian(rng, D, cond)
    mu = 0.3 * rng.standard_normal(D) if with_mean else np.zeros(D)
    tgt = O.Gaussian(Lam, mu)
    model = DensityModel(dim=D, target=GaussianTarget(Lam, mu if with_mean else None))
    return rng, tgt, model


@pytest.mark.parametrize("D,T,prec", [(3, 9, "f64"), (17, 40, "f64"), (64, 200, "f32"), (100, 96, "f32"),
                                      (256, 40, "f32"), (64, 200, "bf16_jac")])
def test_gaussian_fwd_and_jvp(D, T, prec):
    from ParallelMCMC import MALA

    rng, tgt, model = _setup(D + T, D)
    eps = 0.5 / np.max(np.linalg.eigvalsh(tgt.Lam))
    X = tgt.mean[:, None] + 0.3 * rng.standard_normal((D, T))
    Xi = rng.standard_normal((D, T))
    Z = np.where(rng.random((D, T)) < 0.5, -1.0, 1.0)
    U = rng.uniform(0.01, 0.99, T)
    dt = np.float64 if prec == "f64" else np.float32
    X, Xi, Z, U = (a.astype(dt) for a in (X, Xi, Z, U))
    ref = O.mala_step_batched_fwd_and_jvp(tgt, X.astype(np.float64), eps, Xi.astype(np.float64),
                                          U.astype(np.float64), Z.astype(np.float64))
    FT, Jt, acc, la = MALA.mala_step_batched_fwd_and_jvp(model, X, eps, Xi, U, Z, precision=prec,
                                                         return_logalpha=True)
    la = la.cpu().numpy()
    if prec == "f64":
        assert np.allclose(la, ref.logalpha, rtol=1e-10, atol=1e-10)
        assert np.array_equal(acc.cpu().numpy().astype(bool), ref.acc)
        assert np.allclose(FT.cpu().numpy(), ref.FT, rtol=1e-12, atol=1e-12)
        assert np.allclose(Jt.cpu().numpy(), ref.Jt, rtol=1e-9, atol=1e-9)
    else:
        assert np.max(np.abs(la - ref.logalpha)) < 1e-3
        clear = np.abs(ref.logalpha - np.log(U.astype(np.float64))) > 1e-2
        assert np.array_equal(acc.cpu().numpy().astype(bool)[clear], ref.acc[clear])
        assert np.max(np.abs(FT.cpu().numpy() - ref.FT)) / np.max(np.abs(ref.FT)) < 1e-5
        tol = 1e-3 if prec == "f32" else 3e-2
        assert np.max(np.abs(Jt.cpu().numpy() - ref.Jt)) / np.max(np.abs(ref.Jt)) < tol


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_gaussian_logdensity(prec):
    from ParallelMCMC import MALA

    rng, tgt, model = _setup(5, 40)
    S = tgt.mean[:, None] + rng.standard_normal((40, 50))
    dt = np.float64 if prec == "f64" else np.float32
    lp = MALA.logdensity_batch(model, S.astype(dt), precision=prec).cpu().numpy()
    assert np.allclose(lp, tgt.logp_batch(S.astype(dt).astype(np.float64)), rtol=1e-12, atol=1e-9)


def test_gaussian_deer_solve_f64_recovers_sequential():
    from ParallelMCMC import DEER

    rng, tgt, model = _setup(12, 8, cond=1.0)
    D, T = 8, 64
    eps = 0.3 / np.max(np.linalg.eigvalsh(tgt.Lam))
    x0 = tgt.mean + rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    _, info_o = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=400, damping=1.0, seed=2)
    rec = DEER.TapedRecursion(model, eps, Xi, U)
    S, info = DEER.solve(rec, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=400, damping=1.0, seed=2,
                         return_info=True)
    assert info.converged
    assert info.iters <= info_o.iters + 1
    assert np.max(np.abs(S.cpu().numpy() - ref)) < 1e-10
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


@pytest.mark.parametrize("prec", ["f32", "bf16_jac"])
def test_gaussian_deer_solve_f32_d64(prec):
    """D=64 correlated Gaussian (condition number 10), T=256, eps*lambda_max=0.3:
    the GPU matches the sequential trajectory at the 1e-5 relative bar with
    identical decisions (the FP64 oracle takes 55 :stoch_diag iterations).
    At larger eps*lambda_max the Hutchinson estimate's off-diagonal error
    amplifies FP32 rounding between iterations and FP32 :stoch_diag stalls
    above 1e-6 tolerances (:full converges there)."""
    from ParallelMCMC import DEER

    rng, tgt, model = _setup(13, 64, cond=1.0)
    D, T = 64, 256
    eps = 0.3 / np.max(np.linalg.eigvalsh(tgt.Lam))
    x0 = (tgt.mean + rng.standard_normal(D)).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                                  U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision=prec)
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=400, damping=1.0, seed=4, return_info=True)
    assert info.converged, info
    assert info.iters <= T + 1
    assert O.trace_error(S.cpu().numpy().astype(np.float64), ref) < 1e-5
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)
