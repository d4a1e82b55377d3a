"""GPU parity of jacobian=:diag (the exact Jacobian diagonal) on the GEMM
targets, logistic regression and dense Gaussian: the surrogate's log-alpha
gradient from two Hessian-vector GEMM passes plus diag(H_x), against the
oracle's closed form (mala_jac_diag_batched, itself pinned against the
reference's basis-vector JVPs, src/DEER/DEER.jl:327-342).

* one Newton step (FP64) vs the oracle's step at 1e-12, for :diag and, with
  the same probe stream, :stoch_diag (the tape's uniforms enter log u in
  the iterate's precision);
* FP64 solves vs sequential MALA at 1e-10 with every decision identical and
  iteration counts no worse than the oracle's :diag DEER (+1);
* FP32 / BF16-Jacobian solves vs sequential MALA at the 1e-5 relative bar
  with identical decisions.
"""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _logistic(seed, N, D, prior_scale=1.0):
    from ParallelMCMC import DensityModel, LogisticTarget

    rng = np.random.default_rng(seed)
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    return rng, O.LogReg(Xd, y, prior_scale), DensityModel(dim=D, target=LogisticTarget(Xd, y, prior_scale))


def _gauss(seed, D, cond=2.0, with_mean=True):
    from ParallelMCMC import DensityModel, GaussianTarget

    rng = np.random.default_rng(seed)
    Lam, ev = O.make_corr_gaussian(rng, D, cond)
    mu = 0.3 * rng.standard_normal(D) if with_mean else np.zeros(D)
    model = DensityModel(dim=D, target=GaussianTarget(Lam, mu if with_mean else None))
    return rng, O.Gaussian(Lam, mu), model, ev


@pytest.mark.parametrize("jac", ["diag", "stoch_diag"])
@pytest.mark.parametrize("kind,D,T,damping", [("logistic", 7, 50, 1.0), ("logistic", 40, 300, 0.6),
                                              ("gauss", 9, 70, 1.0), ("gauss", 64, 260, 0.8),
                                              ("gauss", 200, 40, 1.0)])
def test_diag_newton_step_f64_matches_oracle(kind, D, T, damping, jac):
    from ParallelMCMC import DEER

    if kind == "logistic":
        rng, tgt, model = _logistic(300 + D, 400, D, prior_scale=0.9)
        eps = 0.02
        x0 = 0.3 * rng.standard_normal(D)
        S_in = 0.3 * rng.standard_normal((D, T))
    else:
        rng, tgt, model, ev = _gauss(300 + D, D)
        eps = 0.8 / ev.max()
        x0 = tgt.mean + rng.standard_normal(D)
        S_in = tgt.mean[:, None] + 0.5 * rng.standard_normal((D, T))
    Xi, U = O.make_tape(rng, D, T)
    ref, acc_ref, _ = O.deer_update(tgt, eps, Xi, U, x0, S_in, jacobian=jac, damping=damping, seed=5)
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision="f64")
    S, info = DEER.solve(rec, x0, init=S_in, tol_abs=0.0, tol_rel=0.0, maxiter=1, jacobian=jac,
                         damping=damping, seed=5, return_info=True, copy_result=True)
    S = S.cpu().numpy()
    err = np.max(np.abs(S - ref))
    print(f"{kind} D={D} T={T}: max|S-ref| = {err:.3e}", flush=True)
    assert err <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


@pytest.mark.parametrize("kind", ["logistic", "gauss"])
def test_diag_deer_f64_recovers_sequential(kind):
    from ParallelMCMC import DEER

    if kind == "logistic":
        rng, tgt, model = _logistic(22, 300, 8)
        D, T, eps = 8, 120, 0.02
        x0 = 0.3 * rng.standard_normal(D)
    else:
        rng, tgt, model, ev = _gauss(23, 12, cond=2.0)
        D, T = 12, 150
        eps = 0.9 / ev.max()
        x0 = tgt.mean + rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    _, info_o = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=400, jacobian="diag",
                             damping=1.0)
    rec = DEER.TapedRecursion(model, eps, Xi, U)
    S, info = DEER.solve(rec, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=400, jacobian="diag", damping=1.0,
                         return_info=True)
    print(f"{kind}: iters gpu {info.iters} oracle {info_o.iters}", flush=True)
    assert info.converged, info
    assert info.iters <= info_o.iters + 1
    assert np.max(np.abs(S.cpu().numpy() - ref)) < 1e-10
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


@pytest.mark.parametrize("kind", ["logistic", "gauss"])
@pytest.mark.parametrize("prec", ["f32", "bf16_jac"])
def test_diag_deer_f32_bf16(kind, prec):
    """Reduced C1 / dense-Gaussian shapes: FP32 states (BF16 Jacobian
    products) within 1e-5 relative of sequential MALA, decisions identical."""
    from ParallelMCMC import DEER

    if kind == "logistic":
        rng, tgt, model = _logistic(7, 2000, 100)
        D, T, eps = 100, 512, 5e-3
        x0 = (0.1 * rng.standard_normal(D)).astype(np.float32)
    else:
        rng, tgt, model, ev = _gauss(31, 64, cond=3.0)
        D, T = 64, 600
        eps = 1.0 / ev.max()
        x0 = (tgt.mean + rng.standard_normal(D)).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                                  U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision=prec)
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=T + 1, jacobian="diag", damping=1.0,
                         return_info=True)
    err = O.trace_error(S.cpu().numpy().astype(np.float64), ref)
    print(f"{kind} {prec}: iters {info.iters}, trace error {err:.3e}", flush=True)
    assert info.converged, info
    assert err < 1e-5
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


def test_diag_gemm_rejects_wide_targets():
    """D > 256 on a GEMM target: a clear NotImplementedError, not a fallback."""
    from ParallelMCMC import DEER

    rng, tgt, model, ev = _gauss(5, 260)
    D, T = 260, 8
    Xi, U = O.make_tape(rng, D, T)
    rec = DEER.TapedRecursion(model, 0.5 / ev.max(), Xi, U, precision="f64")
    with pytest.raises(NotImplementedError):
        DEER.solve(rec, tgt.mean, maxiter=2, jacobian="diag")
