"""GPU parity of the adjoint (reverse-time DEER over the transposed
Jacobians, accept decisions held fixed) against the oracle's sequential
adjoint (pinned by finite differences in tests/test_oracle_adjoint.py):
gradients of L = mean_t |x_t|^2 with respect to eps, the logistic prior scale
and x0, for the three GPU targets, in FP64 (1e-9) and FP32 (1e-4 relative),
and through the windowed, checkpointed autograd Function."""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _solve_and_adjoint(model, tgt, eps, Xi, U, x0, prec, tol, jacobian="stoch_diag"):
    from ParallelMCMC import DEER

    rec = DEER.TapedRecursion(model, eps, Xi, U, precision=prec)
    S, info = DEER.solve(rec, x0, tol_abs=tol, tol_rel=tol, maxiter=500, damping=1.0, return_info=True,
                         copy_result=True, jacobian=jacobian)
    assert info.converged, info
    Sn = S.cpu().numpy().astype(np.float64)
    acc = info.accepted.cpu().numpy().astype(bool)
    G = 2.0 * Sn / Sn.shape[1]
    de, dh, dx0, ainfo = DEER.adjoint(rec, x0, S, info.accepted, G.astype(Xi.dtype))
    assert ainfo.converged, ainfo
    ref = O.mala_adjoint(tgt, eps, Xi.astype(np.float64), np.asarray(x0, np.float64), Sn, acc, G)
    print(f"{prec}: adjoint iters {ainfo.iters}; deps {de:.10e} vs {ref[0]:.10e}; dh {dh:.6e} vs {ref[1]:.6e}",
          flush=True)
    return (de, dh, dx0.cpu().numpy()), ref


def _close(a, b, rel):
    return abs(a - b) <= rel * max(abs(b), 1e-30)


def test_adjoint_diag_gaussian_f64():
    from ParallelMCMC import DensityModel, DiagGaussianTarget

    rng = np.random.default_rng(41)
    D, T, eps = 5, 300, 0.2
    prec, mean = rng.uniform(0.5, 2.0, D), rng.standard_normal(D)
    tgt = O.DiagGaussian(prec, mean)
    model = DensityModel(dim=D, target=DiagGaussianTarget(prec, mean))
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    (de, dh, dx0), ref = _solve_and_adjoint(model, tgt, eps, Xi, U, x0, "f64", 1e-13)
    assert _close(de, ref[0], 1e-10) and dh == 0.0
    assert np.allclose(dx0, ref[2], rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("scale", [0.5, 2.0])
def test_adjoint_logistic_f64(scale):
    from ParallelMCMC import DensityModel, LogisticTarget

    rng = np.random.default_rng(42)
    N, D, T, eps = 400, 6, 200, 0.02
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    tgt = O.LogReg(Xd, y, scale)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y, scale))
    x0 = 0.3 * rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    (de, dh, dx0), ref = _solve_and_adjoint(model, tgt, eps, Xi, U, x0, "f64", 1e-13)
    assert _close(de, ref[0], 1e-9)
    assert _close(dh, ref[1], 1e-9)
    assert np.allclose(dx0, ref[2], rtol=1e-9, atol=1e-13)


def test_adjoint_logistic_f32_mfma():
    from ParallelMCMC import DensityModel, LogisticTarget

    rng = np.random.default_rng(43)
    N, D, T, eps = 2000, 20, 1024, 5e-3
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    Xd = Xd.astype(np.float32).astype(np.float64)
    tgt = O.LogReg(Xd, y, 1.5)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y, 1.5))
    x0 = (0.3 * rng.standard_normal(D)).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    (de, dh, dx0), ref = _solve_and_adjoint(model, tgt, eps, Xi, U, x0, "f32", 1e-6)
    assert _close(de, ref[0], 1e-4)
    assert _close(dh, ref[1], 1e-4)
    assert np.max(np.abs(dx0 - ref[2])) <= 1e-4 * np.max(np.abs(ref[2]))


def test_adjoint_dense_gaussian_f32():
    from ParallelMCMC import DensityModel, GaussianTarget

    rng = np.random.default_rng(44)
    D, T = 32, 512
    Lam, ev = O.make_corr_gaussian(rng, D, 1.0)
    Lam = Lam.astype(np.float32).astype(np.float64)
    tgt = O.Gaussian(Lam, np.zeros(D))
    model = DensityModel(dim=D, target=GaussianTarget(Lam))
    eps = 0.2 / ev.max()
    x0 = rng.standard_normal(D).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    # the forward solve uses the dense Jacobian (diagonal estimates converge
    # slowly on correlated Gaussians); the adjoint sweeps use Hutchinson's diagonal
    (de, dh, dx0), ref = _solve_and_adjoint(model, tgt, eps, Xi, U, x0, "f32", 1e-6, jacobian="full")
    assert _close(de, ref[0], 1e-4) and dh == 0.0
    assert np.max(np.abs(dx0 - ref[2])) <= 1e-4 * np.max(np.abs(ref[2]))


def test_windowed_autograd_matches_single_window_and_oracle():
    """T=1000 in windows of 256 (checkpoint re-solves, carried dL/dx across
    window boundaries) through torch.autograd: same gradients as the
    oracle's adjoint over the whole trajectory."""
    import torch

    from ParallelMCMC import DensityModel, LogisticTarget
    from ParallelMCMC.Adjoint import DEERTrajectory

    rng = np.random.default_rng(45)
    N, D, T, eps, scale = 300, 5, 1000, 0.02, 0.8
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    tgt = O.LogReg(Xd, y, scale)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y, scale))
    x0n = 0.3 * rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    eps_t = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    s_t = torch.tensor(scale, dtype=torch.float64, requires_grad=True)
    x0 = torch.tensor(x0n, dtype=torch.float64, device="cuda", requires_grad=True)
    opts = {"precision": "f64", "solve": dict(tol_abs=1e-13, tol_rel=1e-13, maxiter=500, damping=1.0)}
    S = DEERTrajectory.apply(eps_t, s_t, x0, model, Xi, U, 256, opts)
    L = (S * S).sum() / T
    L.backward()
    ref_S, acc, _ = O.run_mala_sequential_taped(tgt, x0n, eps, Xi, U)
    assert np.max(np.abs(S.detach().cpu().numpy() - ref_S)) < 1e-10
    ref = O.mala_adjoint(tgt, eps, Xi, x0n, ref_S, acc, 2.0 * ref_S / T)
    print(f"windowed: deps {eps_t.grad.item():.10e} vs {ref[0]:.10e}", flush=True)
    assert _close(eps_t.grad.item(), ref[0], 1e-9)
    assert _close(s_t.grad.item(), ref[1], 1e-9)
    assert np.allclose(x0.grad.cpu().numpy(), ref[2], rtol=1e-9, atol=1e-13)
