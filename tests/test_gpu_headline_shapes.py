"""GPU parity on the headline shapes (the batched and single-chain workloads
the bench measures), against the FP64 sequential MALA oracle:

* batched chains of the uneven-chain workload's classes in one plan with a
  4096-column sliding window: dense Gaussians at D = 100, 200, 256 with
  precision condition 1e3 (eps = 0.5 / lambda_max; these crawl about one
  column per Newton iteration under :stoch_diag), and logistic regressions
  at N = 2000 with D = 150 and D = 256 (the widest GEMM class) and T longer
  than the window; BF16 Jacobians and FP32;
* one full logistic-regression block at N = 10^4, D = 100, T = 4096,
  eps = 5e-3 (config 1 of the reference's benchmark), FP32 and BF16.

Bar: trace error <= 1e-5 relative and every accept decision identical.
"""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _headline_chains(seed):
    from ParallelMCMC import DensityModel, GaussianTarget, LogisticTarget
    from ParallelMCMC.Batched import Chain

    rng = np.random.default_rng(seed)
    out = []
    for D, T in [(100, 3000), (200, 2500), (256, 2000)]:
        Lam, ev = O.make_corr_gaussian(rng, D, 3.0)
        Lam = Lam.astype(np.float32).astype(np.float64)
        ev = np.linalg.eigvalsh(Lam)
        tgt = O.Gaussian(Lam, np.zeros(D))
        eps = 0.5 / ev.max()
        x0 = (rng.standard_normal(D) / np.sqrt(ev.mean())).astype(np.float32)
        Xi, U = O.make_tape(rng, D, T, np.float32)
        out.append((tgt, Chain(DensityModel(dim=D, target=GaussianTarget(Lam)), eps, x0, Xi, U)))
    for D, T in [(150, 5000), (256, 4500)]:
        Xd, y, _ = O.make_logreg_data(rng, 2000, D)
        Xd = Xd.astype(np.float32).astype(np.float64)
        tgt = O.LogReg(Xd, y, 1.0)
        eps = 0.15 / np.linalg.eigvalsh(0.25 * Xd.T @ Xd + np.eye(D)).max()
        x0 = (0.1 * rng.standard_normal(D)).astype(np.float32)
        Xi, U = O.make_tape(rng, D, T, np.float32)
        out.append((tgt, Chain(DensityModel(dim=D, target=LogisticTarget(Xd, y, 1.0)), eps, x0, Xi, U)))
    return out


@pytest.mark.parametrize("prec", ["bf16_jac", "f32"])
def test_batched_headline_shapes(prec):
    from ParallelMCMC.Batched import BatchPlan

    pairs = _headline_chains(2024)
    W = 4096
    plan = BatchPlan([c for _, c in pairs], window=W, precision=prec, tol_abs=1e-6, tol_rel=1e-5, maxiter=W + 1)
    for (tgt, ch), (S, acc, info) in zip(pairs, plan.solve()):
        assert info.converged, info
        ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, ch.x0.astype(np.float64), ch.eps,
                                                      ch.Xi.astype(np.float64), ch.U.astype(np.float64))
        err = O.trace_error(S.cpu().numpy().astype(np.float64), ref)
        mism = int(np.sum(acc.cpu().numpy().astype(bool) != acc_ref))
        print(f"{prec} {type(tgt).__name__} D={tgt.dim} T={ch.Xi.shape[1]}: iters {info.iters}, "
              f"trace error {err:.2e}, mismatches {mism}", flush=True)
        assert err <= 1e-5
        assert mism == 0


@pytest.mark.parametrize("prec", ["f32", "bf16_jac"])
def test_full_c1_block(prec):
    from ParallelMCMC import DEER, DensityModel, LogisticTarget

    rng = np.random.default_rng(1)
    N, D, T, eps = 10_000, 100, 4096, 5e-3
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    Xd = Xd.astype(np.float32).astype(np.float64)
    tgt = O.LogReg(Xd, y, 1.0)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y, 1.0))
    x0 = rng.standard_normal(D).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                                  U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision=prec)
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-5, maxiter=T + 1, damping=1.0, return_info=True)
    err = O.trace_error(S.cpu().numpy().astype(np.float64), ref)
    mism = int(np.sum(info.accepted.cpu().numpy().astype(bool) != acc_ref))
    print(f"C1 {prec}: iters {info.iters}, trace error {err:.2e}, mismatches {mism}", flush=True)
    assert info.converged, info
    assert err <= 1e-5
    assert mism == 0
