"""GPU parity of jacobian=:full (dense Gaussian target): structured Jacobian
products on MFMA (FP32 / BF16) or VALU (FP64), tile matrix-scan, correction
form of the Newton update.

* one Newton step vs the oracle's dense step (deer_update with the exact
  dense Jacobian): FP64 at 1e-10, several tiles with a ragged last one;
* solves vs the sequential MALA oracle: FP64 trajectory within 1e-10,
  FP32 / BF16 within 1e-5 relative; accept decisions identical; iteration
  counts no worse than the oracle's full-Jacobian DEER (+1) and far below
  the diagonal estimators'.
"""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _setup(seed, D, cond=2.0, with_mean=True):
    from ParallelMCMC import DensityModel, GaussianTarget

    rng = np.random.default_rng(seed)
    Lam, ev = O.make_corr_gaussian(rng, D, cond)
    mu = 0.3 * rng.standard_normal(D) if with_mean else np.zeros(D)
    tgt = O.Gaussian(Lam, mu)
    model = DensityModel(dim=D, target=GaussianTarget(Lam, mu if with_mean else None))
    return rng, tgt, model, ev


@pytest.mark.parametrize("jac", ["full", "full_sigmoid"])
# T = 4477 (70 tiles) and 16384 / 16320 (256 / 255 tiles: 32 groups of 8 maps)
# take the one-launch scan over the tile maps; 38417 (601 tiles) the recursive
# form through three reduce levels of 8-map tiles
@pytest.mark.parametrize("D,T,damping", [(5, 200, 0.7), (16, 64, 1.0), (64, 130, 1.0), (16, 4477, 1.0),
                                         (4, 16384, 1.0), (6, 16320, 0.8), (4, 38417, 0.9)])
def test_full_newton_step_f64_matches_dense_oracle(D, T, damping, jac):
    from ParallelMCMC import DEER

    rng, tgt, model, ev = _setup(100 + D, D)
    eps = 0.8 / ev.max()
    x0 = tgt.mean + rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    S_in = tgt.mean[:, None] + 0.5 * rng.standard_normal((D, T))
    ref, acc_ref, _ = O.deer_update(tgt, eps, Xi, U, x0, S_in, jacobian=jac, damping=damping)
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision="f64")
    S, info = DEER.solve(rec, x0, init=S_in, tol_abs=0.0, tol_rel=0.0, maxiter=1, jacobian=jac,
                         damping=damping, return_info=True, copy_result=True)
    S = S.cpu().numpy()
    print(f"D={D} T={T}: max|S-ref| = {np.max(np.abs(S - ref)):.3e}", flush=True)
    assert np.max(np.abs(S - ref)) <= 1e-10 * max(1.0, np.max(np.abs(ref)))
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


@pytest.mark.parametrize("jac", ["full", "full_sigmoid"])
@pytest.mark.parametrize("prec,tol", [("f32", 2e-4), ("bf16_jac", 5e-2)])
@pytest.mark.parametrize("T", [150, 2567])
def test_full_newton_step_f32_bf16(prec, tol, jac, T):
    """One step at D=64 over 3 tiles (41 at T = 2567: one reduce level over the
    tile maps): FP32 products agree with the dense FP64 step to FP32 accuracy;
    BF16 products perturb only the Jacobian part."""
    from ParallelMCMC import DEER

    rng, tgt, model, ev = _setup(7, 64)
    D = 64
    eps = 0.8 / ev.max()
    x0 = tgt.mean + rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    S_in = tgt.mean[:, None] + 0.2 * rng.standard_normal((D, T))
    ref, acc_ref, la = O.deer_update(tgt, eps, Xi, U, x0, S_in, jacobian=jac, damping=1.0)
    rec = DEER.TapedRecursion(model, eps, Xi.astype(np.float32), U.astype(np.float32), precision=prec)
    S, info = DEER.solve(rec, x0.astype(np.float32), init=S_in.astype(np.float32), tol_abs=0.0, tol_rel=0.0,
                         maxiter=1, jacobian=jac, damping=1.0, return_info=True, copy_result=True)
    S = S.cpu().numpy().astype(np.float64)
    step = np.max(np.abs(ref - S_in))
    err = np.max(np.abs(S - ref)) / step
    print(f"{prec}: step {step:.3e}, relative step error {err:.3e}", flush=True)
    assert err < tol
    clear = np.abs(la - np.log(U)) > 1e-3
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool)[clear], acc_ref[clear])


@pytest.mark.parametrize("jac", ["full", "full_sigmoid"])
def test_full_deer_f64_recovers_sequential(jac):
    from ParallelMCMC import DEER

    rng, tgt, model, ev = _setup(21, 16, cond=3.0)
    D, T = 16, 300
    eps = 1.0 / ev.max()
    x0 = tgt.mean + rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    _, info_o = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=200, jacobian=jac,
                             damping=1.0)
    rec = DEER.TapedRecursion(model, eps, Xi, U)
    S, info = DEER.solve(rec, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=200, jacobian=jac, damping=1.0,
                         return_info=True)
    print(f"iters gpu {info.iters} oracle {info_o.iters}", flush=True)
    assert info.converged
    assert info.iters <= info_o.iters + 1
    assert np.max(np.abs(S.cpu().numpy() - ref)) < 1e-10
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


@pytest.mark.parametrize("prec", ["f32", "bf16_jac"])
def test_full_deer_f32_bf16_d64(prec):
    """D=64, condition number 1e3, eps * lambda_max = 1, T=1000 (16 tiles, the
    last ragged), decision-frozen Jacobian: the trajectory matches sequential
    MALA at the 1e-5 relative bar with identical decisions, in far fewer
    iterations than :stoch_diag (the oracle needs 14)."""
    from ParallelMCMC import DEER

    rng, tgt, model, ev = _setup(31, 64, cond=3.0)
    D, T = 64, 1000
    eps = 1.0 / ev.max()
    x0 = (tgt.mean + rng.standard_normal(D)).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                                  U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision=prec)
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=300, jacobian="full", damping=1.0,
                         return_info=True)
    err = O.trace_error(S.cpu().numpy().astype(np.float64), ref)
    print(f"{prec}: iters {info.iters}, trace error {err:.3e}", flush=True)
    assert info.converged, info
    assert err < 1e-5
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)
    _, info_sd = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=300, jacobian="stoch_diag",
                            damping=1.0, return_info=True)
    print(f"{prec}: stoch_diag iters {info_sd.iters}", flush=True)
    assert info.iters * 3 < info_sd.iters



@pytest.mark.parametrize("prec,D,T", [("f64", 16, 2000), ("f32", 64, 4096), ("bf16_jac", 64, 4096)])
def test_full_deer_final_prefix_at_the_stability_edge(prec, D, T):
    """eps * lambda_max = 1.9 on a condition-1e3 precision (eigenvalues
    10^U(0, 3)): low acceptance, so
    Newton advances a few columns per iteration and the solve leaves out the
    tiles whose residual is exactly 0 (the final prefix).  The result is still
    sequential MALA: FP64 within 1e-10, FP32 / BF16 within 1e-5 relative,
    every decision identical."""
    from ParallelMCMC import DEER

    rng, tgt, model, ev = _setup(41 + D, D, cond=3.0, with_mean=False)
    eps = 1.9 / ev.max()
    x0 = rng.standard_normal(D)
    dt = np.float64 if prec == "f64" else np.float32
    Xi, U = O.make_tape(rng, D, T, dt)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi.astype(np.float64), U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision=prec)
    tol = (1e-12, 1e-11) if prec == "f64" else (1e-6, 1e-6)
    S, info = DEER.solve(rec, x0.astype(dt), tol_abs=tol[0], tol_rel=tol[1], maxiter=T + 1, jacobian="full",
                         damping=1.0, return_info=True)
    err = O.trace_error(S.cpu().numpy().astype(np.float64), ref)
    print(f"{prec} D={D} T={T}: iters {info.iters}, acceptance {acc_ref.mean():.2f}, trace error {err:.3e}",
          flush=True)
    assert info.converged, info
    assert err < (1e-10 if prec == "f64" else 1e-5)
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


@pytest.mark.parametrize("prec", ["f32", "bf16_jac"])
def test_full_deer_multi_iteration_32_groups(prec):
    """T = 16384 at D = 8 (256 tiles: the one-launch scan over 32 groups of 8
    tile maps), solved to convergence in FP32 / BF16: the spine's epochs are
    reused over many launches while re-tiling drops the final prefix and the
    group count shrinks.  Sequential MALA within 1e-5 relative, every
    decision identical."""
    from ParallelMCMC import DEER

    rng, tgt, model, ev = _setup(57, 8, cond=2.0)
    D, T = 8, 16384
    eps = 1.2 / ev.max()
    x0 = (tgt.mean + rng.standard_normal(D)).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                                  U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision=prec)
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=400, jacobian="full", damping=1.0,
                         return_info=True)
    err = O.trace_error(S.cpu().numpy().astype(np.float64), ref)
    print(f"{prec} D={D} T={T}: iters {info.iters}, acceptance {acc_ref.mean():.2f}, trace error {err:.3e}",
          flush=True)
    assert info.converged, info
    assert info.iters > 2
    assert err < 1e-5
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)
