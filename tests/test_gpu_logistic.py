"""GPU parity of the logistic-regression (GEMM) target path against the
oracle: the batched MALA step + JVP through the fused GEMM passes, trajectory
log-densities, and the full DEER solve (FP64 plain path at the 1e-10 bar,
FP32 MFMA path at the 1e-5 relative bar) — the logistic model of
benchmarks/ParallelMCMCBenchmarks/src/models/bayes_logreg.jl and
test/test-DEER-Turing-Logistic.jl of the reference."""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _setup(seed, N, D, prior_scale=1.0):
    from ParallelMCMC import DensityModel, LogisticTarget

    rng = np.random.default_rng(seed)
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    tgt = O.LogReg(Xd, y, prior_scale)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y, prior_scale))
    return rng, tgt, model


@pytest.mark.parametrize("N,D,T,prec", [(50, 3, 9, "f64"), (333, 7, 40, "f64"), (1000, 100, 96, "f32"),
                                        (77, 17, 33, "f32"), (2000, 200, 40, "f32")])
def test_logistic_fwd_and_jvp(N, D, T, prec):
    from ParallelMCMC import MALA

    rng, tgt, model = _setup(N + D + T, N, D, prior_scale=0.8)
    X = 0.3 * rng.standard_normal((D, T))
    Xi = rng.standard_normal((D, T))
    Z = np.where(rng.random((D, T)) < 0.5, -1.0, 1.0)
    U = rng.uniform(0.01, 0.99, T)
    eps = 0.01
    dt = np.float64 if prec == "f64" else np.float32
    X, Xi, Z, U = (a.astype(dt) for a in (X, Xi, Z, U))
    ref = O.mala_step_batched_fwd_and_jvp(tgt, X.astype(np.float64), eps, Xi.astype(np.float64),
                                          U.astype(np.float64), Z.astype(np.float64))
    FT, Jt, acc, la = MALA.mala_step_batched_fwd_and_jvp(model, X, eps, Xi, U, Z, precision=prec, return_logalpha=True)
    la = la.cpu().numpy()
    if prec == "f64":
        assert np.allclose(la, ref.logalpha, rtol=1e-10, atol=1e-10)
        assert np.array_equal(acc.cpu().numpy().astype(bool), ref.acc)
        assert np.allclose(FT.cpu().numpy(), ref.FT, rtol=1e-12, atol=1e-12)
        assert np.allclose(Jt.cpu().numpy(), ref.Jt, rtol=1e-9, atol=1e-9)
    else:
        # FP32 GEMMs: logalpha to ~1e-4 absolute; decisions identical away from ties
        assert np.max(np.abs(la - ref.logalpha)) < 2e-3
        clear = np.abs(ref.logalpha - np.log(U.astype(np.float64))) > 1e-2
        assert np.array_equal(acc.cpu().numpy().astype(bool)[clear], ref.acc[clear])
        assert np.max(np.abs(FT.cpu().numpy() - ref.FT)) / np.max(np.abs(ref.FT)) < 1e-5
        assert np.max(np.abs(Jt.cpu().numpy() - ref.Jt)) / np.max(np.abs(ref.Jt)) < 1e-3


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_logistic_logdensity_batch(prec):
    from ParallelMCMC import MALA

    rng, tgt, model = _setup(4, 500, 12, 1.3)
    S = 0.5 * rng.standard_normal((12, 70))
    dt = np.float64 if prec == "f64" else np.float32
    lp = MALA.logdensity_batch(model, S.astype(dt), precision=prec).cpu().numpy()
    ref = tgt.logp_batch(S.astype(dt).astype(np.float64))
    assert np.allclose(lp, ref, rtol=1e-12, atol=1e-9)


def test_logistic_deer_solve_f64_recovers_sequential():
    from ParallelMCMC import DEER

    rng, tgt, model = _setup(22, 200, 6)
    D, T, eps = 6, 80, 0.02
    x0 = 0.3 * rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    _, info_o = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=300,
                             damping=0.5, seed=3)
    rec = DEER.TapedRecursion(model, eps, Xi, U)
    S, info = DEER.solve(rec, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=300, damping=0.5, seed=3,
                         return_info=True)
    assert info.converged
    assert info.iters <= info_o.iters + 1
    assert np.max(np.abs(S.cpu().numpy() - ref)) < 1e-10
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


def test_logistic_deer_solve_f32_mini_c1():
    """Logistic workload at reduced size (N=2000, D=100, T=512, eps=5e-3),
    FP32 MFMA path vs the FP64 sequential oracle on the same float32 tape."""
    from ParallelMCMC import DEER

    rng, tgt, model = _setup(7, 2000, 100)
    D, T, eps = 100, 512, 5e-3
    x0 = (0.1 * rng.standard_normal(D)).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    ref, acc_ref, la_ref = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                                       U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision="f32")
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=500, damping=1.0, seed=11,
                         return_info=True)
    assert info.converged, info
    err = O.trace_error(S.cpu().numpy().astype(np.float64), ref)
    assert err < 1e-5, err
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


def test_logistic_bf16_jacobian_keeps_residual_path_fp32():
    """BF16 Jacobian products change Jt only: FT, decisions and log-alpha are
    bit-identical to the FP32 path, Jt agrees to BF16 accuracy."""
    from ParallelMCMC import MALA

    rng, tgt, model = _setup(31, 3000, 100)
    D, T, eps = 100, 256, 5e-3
    X = (0.2 * rng.standard_normal((D, T))).astype(np.float32)
    Xi = rng.standard_normal((D, T)).astype(np.float32)
    Z = np.where(rng.random((D, T)) < 0.5, -1.0, 1.0).astype(np.float32)
    U = rng.uniform(0.01, 0.99, T).astype(np.float32)
    F32 = MALA.mala_step_batched_fwd_and_jvp(model, X, eps, Xi, U, Z, precision="f32", return_logalpha=True)
    BF = MALA.mala_step_batched_fwd_and_jvp(model, X, eps, Xi, U, Z, precision="bf16_jac", return_logalpha=True)
    assert np.array_equal(F32[0].cpu().numpy(), BF[0].cpu().numpy())
    assert np.array_equal(F32[2].cpu().numpy(), BF[2].cpu().numpy())
    assert np.array_equal(F32[3].cpu().numpy(), BF[3].cpu().numpy())
    j32, jbf = F32[1].cpu().numpy(), BF[1].cpu().numpy()
    assert np.max(np.abs(jbf - j32)) / np.max(np.abs(j32)) < 2e-2
    ref = O.mala_step_batched_fwd_and_jvp(tgt, X.astype(np.float64), eps, Xi.astype(np.float64),
                                          U.astype(np.float64), Z.astype(np.float64))
    assert np.max(np.abs(jbf - ref.Jt)) / np.max(np.abs(ref.Jt)) < 2e-2


def test_logistic_deer_solve_bf16_jac_mini_c1():
    """Same fixed point as FP32 (trace error vs the FP64 sequential run
    <= 1e-5, identical decisions); only the iteration count may change."""
    from ParallelMCMC import DEER

    rng, tgt, model = _setup(7, 2000, 100)
    D, T, eps = 100, 512, 5e-3
    x0 = (0.1 * rng.standard_normal(D)).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                                  U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision="bf16_jac")
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=500, damping=1.0, seed=11,
                         return_info=True)
    assert info.converged, info
    assert O.trace_error(S.cpu().numpy().astype(np.float64), ref) < 1e-5
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


def test_fp64_decision_refinement_near_threshold(monkeypatch):
    """Columns whose log alpha sits within 1e-7 of log u: the FP32 products
    cannot decide them (their rounding moves log alpha by ~1e-5), the FP64
    recompute must reproduce the FP64 decisions exactly.  Negative control:
    with the recompute off, some of them flip."""
    from ParallelMCMC import DEER, DensityModel, LogisticTarget

    rng = np.random.default_rng(77)
    N, D, T, eps = 2000, 20, 512, 1e-2
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    tgt = O.LogReg(Xd, y)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y))
    # states from the chain itself (burned in), so log alpha spreads around 0
    Xb, Ub = O.make_tape(rng, D, 1000)
    Sb, _, _ = O.run_mala_sequential_taped(tgt, np.zeros(D), eps, Xb, Ub)
    x0 = Sb[:, -1].astype(np.float32)
    Xi = rng.standard_normal((D, T)).astype(np.float32)
    S_in, _, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64), rng.random(T))
    S_in = S_in.astype(np.float32)
    Xbar = O._xbar(x0.astype(np.float64), S_in.astype(np.float64))
    la = O.mala_step_batched_fwd_and_jvp(tgt, Xbar, eps, Xi.astype(np.float64), np.full(T, 0.5),
                                         np.ones((D, T))).logalpha
    near = la < -1e-5
    delta = np.where(np.arange(T) % 2 == 0, 1e-7, -1e-7) * (1.0 + np.abs(la))
    U = np.where(near, np.exp(la - delta), rng.uniform(0.05, 0.95, T))
    U32 = U.astype(np.float32)
    lu = np.log(U32.astype(np.float64))
    want = lu < la
    assert near.sum() > 50
    rec = DEER.TapedRecursion(model, eps, Xi, U32, precision="f32")

    def decisions():
        _, info = DEER.solve(rec, x0, init=S_in, tol_abs=0.0, tol_rel=0.0, maxiter=1, damping=1.0,
                             return_info=True, copy_result=True)
        return info.accepted.cpu().numpy().astype(bool)

    # maxiter=1 with tol 0: the gate (previous update within 100 tolerance
    # units) never opens, so run the recompute from the first iteration
    monkeypatch.setenv("PMCMC_REFINE", "all")
    got = decisions()
    monkeypatch.setenv("PMCMC_REFINE", "0")
    raw = decisions()
    print(f"near-threshold columns {near.sum()}: mismatches refined {np.sum(got != want)}, "
          f"FP32 only {np.sum(raw != want)}", flush=True)
    assert np.array_equal(got, want)
    assert np.sum(raw[near] != want[near]) > 0


def test_fp64_decisions_on_converged_trajectory(monkeypatch):
    """Solve from scratch with the default gated FP64 recompute on a tape whose
    every twelfth column sits 4e-6 from its accept threshold along the FP64
    chain (inside the FP32 products' log-alpha error, up to ~1e-5 here; such
    columns slow DEER itself down, the FP64 oracle needs 73 iterations).  Every
    returned decision must equal the FP64 decision evaluated at the returned
    trajectory, except columns within 2e-6 of the threshold there (the last
    iteration decides at its input iterate, which differs from the returned
    one by FP32 noise, ~1e-7, moving log alpha by < 1e-6).  Negative control:
    the FP32-only solve gets some of them wrong."""
    from ParallelMCMC import DEER, DensityModel, LogisticTarget

    rng = np.random.default_rng(78)
    N, D, T, eps = 2000, 20, 768, 1e-2
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    tgt = O.LogReg(Xd, y)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y))
    Xb, Ub = O.make_tape(rng, D, 1000)
    Sb, _, _ = O.run_mala_sequential_taped(tgt, np.zeros(D), eps, Xb, Ub)
    x0 = Sb[:, -1].astype(np.float32)
    Xi = rng.standard_normal((D, T)).astype(np.float32)
    U = rng.uniform(0.05, 0.95, T)
    x = x0.astype(np.float64)
    for t in range(T):  # the FP64 chain, placing u near the threshold as it goes
        yv, _ = O.mala_proposal(tgt, x, eps, Xi[:, t].astype(np.float64))
        la = O.mala_logalpha(tgt, x, yv, eps)
        if t % 12 == 0 and la < -1e-4:
            U[t] = np.exp(la - (4e-6 if t % 24 == 0 else -4e-6))
        if np.log(U[t]) < la:
            x = yv
    U32 = U.astype(np.float32)
    lu = np.log(U32.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U32, precision="f32")

    def run(mode):
        if mode is None:
            monkeypatch.delenv("PMCMC_REFINE", raising=False)
        else:
            monkeypatch.setenv("PMCMC_REFINE", mode)
        S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=500, damping=1.0, seed=3,
                             return_info=True, copy_result=True)
        assert info.converged, f"PMCMC_REFINE={mode}: iters {info.iters}, metric {info.metric}"
        S = S.cpu().numpy().astype(np.float64)
        la = O.mala_step_batched_fwd_and_jvp(tgt, O._xbar(x0.astype(np.float64), S), eps,
                                             Xi.astype(np.float64), np.full(T, 0.5), np.ones((D, T))).logalpha
        margin = np.abs(la - lu)
        got = info.accepted.cpu().numpy().astype(bool)
        checked = margin > 2e-6
        return np.sum(got[checked] != (lu < la)[checked]), np.sum(checked & (margin < 2e-5)), info.iters

    bad, close, iters = run(None)
    bad_raw, close_raw, iters_raw = run("0")
    print(f"columns within 2e-5 of the threshold: {close} / {close_raw}; mismatches gated FP64 {bad} "
          f"({iters} iterations), FP32 only {bad_raw} ({iters_raw} iterations)", flush=True)
    assert close >= 25
    assert bad == 0
    assert bad_raw > 0
