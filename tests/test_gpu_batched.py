"""GPU parity of the batched many-chain solve (descriptor-table launches,
window rounds, graph replay): every chain's trajectory equals its sequential
MALA run (1e-5 relative, identical decisions) for chains of uneven D, T and
target kind in one batch; graph replay and direct launches give bit-identical
results; BF16 Jacobians change only the iteration counts."""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _chains(seed):
    from ParallelMCMC import DensityModel, GaussianTarget, LogisticTarget
    from ParallelMCMC.Batched import Chain

    rng = np.random.default_rng(seed)
    out = []
    for D, T, N in [(3, 300, 500), (17, 700, 500), (40, 1000, 800)]:
        Xd, y, _ = O.make_logreg_data(rng, N, D)
        Xd = Xd.astype(np.float32).astype(np.float64)
        tgt = O.LogReg(Xd, y, 1.0)
        ev = np.linalg.eigvalsh(-np.stack([tgt.hvp(np.zeros(D), e) for e in np.eye(D)], axis=1))
        eps = 0.3 / ev.max()
        x0 = (0.2 * rng.standard_normal(D)).astype(np.float32)
        Xi, U = O.make_tape(rng, D, T, np.float32)
        out.append((tgt, Chain(DensityModel(dim=D, target=LogisticTarget(Xd, y, 1.0)), eps, x0, Xi, U)))
    for D, T in [(5, 256, ), (33, 900), (64, 1500)]:
        Lam, ev = O.make_corr_gaussian(rng, D, 1.0)
        Lam = Lam.astype(np.float32).astype(np.float64)
        tgt = O.Gaussian(Lam, np.zeros(D))
        eps = 0.2 / ev.max()
        x0 = rng.standard_normal(D).astype(np.float32)
        Xi, U = O.make_tape(rng, D, T, np.float32)
        out.append((tgt, Chain(DensityModel(dim=D, target=GaussianTarget(Lam)), eps, x0, Xi, U)))
    return out


def _check(pairs, results):
    for (tgt, ch), (S, acc, info) in zip(pairs, results):
        assert info.converged, info
        ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, ch.x0.astype(np.float64), ch.eps,
                                                      ch.Xi.astype(np.float64), ch.U.astype(np.float64))
        err = O.trace_error(S.cpu().numpy().astype(np.float64), ref)
        print(f"D={tgt.dim} T={ch.Xi.shape[1]}: iters {info.iters}, trace error {err:.2e}", flush=True)
        assert err < 1e-5
        assert np.array_equal(acc.cpu().numpy().astype(bool), acc_ref)


@pytest.mark.parametrize("prec", ["f32", "bf16_jac"])
def test_batched_uneven_chains_match_sequential(prec):
    from ParallelMCMC.Batched import BatchPlan

    pairs = _chains(7)
    plan = BatchPlan([c for _, c in pairs], window=256, precision=prec, tol_abs=1e-6, tol_rel=1e-6, maxiter=300)
    _check(pairs, plan.solve())


def _wide_logistic_chains(seed):
    """Four logistic chains at N=2000 with 4096-column windows: once their
    windows have widened to full width they hold 512 live GEMM column blocks,
    and the GEMM passes take the split count sized for 512 workgroups per
    chain (4 splits instead of the plan's 6)."""
    from ParallelMCMC import DensityModel, LogisticTarget
    from ParallelMCMC.Batched import Chain

    rng = np.random.default_rng(seed)
    out = []
    for D, T in [(6, 4200), (12, 4100), (20, 4500), (30, 4096)]:
        Xd, y, _ = O.make_logreg_data(rng, 2000, D)
        Xd = Xd.astype(np.float32).astype(np.float64)
        tgt = O.LogReg(Xd, y, 1.0)
        ev = np.linalg.eigvalsh(-np.stack([tgt.hvp(np.zeros(D), e) for e in np.eye(D)], axis=1))
        eps = 0.15 / ev.max()
        x0 = (0.2 * rng.standard_normal(D)).astype(np.float32)
        Xi, U = O.make_tape(rng, D, T, np.float32)
        out.append((tgt, Chain(DensityModel(dim=D, target=LogisticTarget(Xd, y, 1.0)), eps, x0, Xi, U)))
    return out


def test_batched_many_gemm_blocks_match_sequential():
    from ParallelMCMC.Batched import BatchPlan

    pairs = _wide_logistic_chains(11)
    chains = [c for _, c in pairs]
    a = BatchPlan(chains, window=4096, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-5, maxiter=600).solve()
    _check(pairs, a)
    b = BatchPlan(chains, window=4096, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-5, maxiter=600,
                  use_graph=False).solve()
    for (Sa, acca, ia), (Sb, accb, ib) in zip(a, b):
        assert ia.iters == ib.iters
        assert np.array_equal(Sa.cpu().numpy(), Sb.cpu().numpy())
        assert np.array_equal(acca.cpu().numpy(), accb.cpu().numpy())


def test_batched_graph_and_direct_launches_identical():
    from ParallelMCMC.Batched import BatchPlan

    pairs = _chains(8)
    chains = [c for _, c in pairs]
    a = BatchPlan(chains, window=300, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-6, maxiter=300).solve()
    b = BatchPlan(chains, window=300, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-6, maxiter=300,
                  use_graph=False).solve()
    for (Sa, acca, ia), (Sb, accb, ib) in zip(a, b):
        assert ia.iters == ib.iters
        assert np.array_equal(Sa.cpu().numpy(), Sb.cpu().numpy())
        assert np.array_equal(acca.cpu().numpy(), accb.cpu().numpy())
    # a second solve on the same plan repeats the first
    plan = BatchPlan(chains, window=300, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-6, maxiter=300)
    r1 = [(S.clone(), i.iters) for S, _, i in plan.solve()]
    r2 = plan.solve()
    for (S1, it1), (S2, _, i2) in zip(r1, r2):
        assert it1 == i2.iters and np.array_equal(S1.cpu().numpy(), S2.cpu().numpy())


def test_batched_profiled_solve_counts_pass_flops():
    """A profiled solve (direct launches, events around the GEMM and Lambda
    passes) returns the graph-replayed solve's results bit for bit; with one
    logistic chain whose trajectory fits one window, every live GEMM pass
    covers all N rows and between 1 and T columns, so the counted flops lie
    between 2 x iters x 8 N D and 2 x iters x 8 N D T (launches queued after
    convergence are timed but count no flops)."""
    from ParallelMCMC.Batched import BatchPlan

    pairs = _chains(9)
    chains = [c for _, c in pairs]
    plan = BatchPlan(chains, window=300, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-6, maxiter=300)
    ref = [(S.clone(), a.clone(), i.iters) for S, a, i in plan.solve()]
    plan.set_profiling(True)
    got = plan.solve()
    n_glm, ms_glm, fl_glm = plan.glm_pass_stats()
    n_lam, ms_lam, fl_lam = plan.pass_stats()
    plan.set_profiling(False)
    for (S1, a1, it1), (S2, a2, i2) in zip(ref, got):
        assert it1 == i2.iters
        assert np.array_equal(S1.cpu().numpy(), S2.cpu().numpy())
        assert np.array_equal(a1.cpu().numpy(), a2.cpu().numpy())
    assert n_glm > 0 and ms_glm > 0 and fl_glm > 0
    assert n_lam > 0 and ms_lam > 0 and fl_lam > 0
    # a single logistic chain (D=3, T=300, N=500) in a window of 300 columns
    (tgt, ch), = [pairs[0]]
    one = BatchPlan([ch], window=300, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-6, maxiter=300)
    one.set_profiling(True)
    (_, _, info), = one.solve()
    n1, _, f1 = one.glm_pass_stats()
    D, T, N = tgt.dim, ch.Xi.shape[1], tgt.Xd.shape[0]
    assert n1 >= 2 * info.iters
    assert 2 * info.iters * 8 * N * D <= f1 <= 2 * info.iters * 8 * N * D * T


def _near_threshold_chain(seed, N=2000, D=20, T=768, eps=1e-2):
    """A logistic chain whose tape puts every twelfth column 4e-6 from its
    accept threshold along the FP64 chain (inside the FP32 products' error)."""
    from ParallelMCMC import DensityModel, LogisticTarget
    from ParallelMCMC.Batched import Chain

    rng = np.random.default_rng(seed)
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    tgt = O.LogReg(Xd, y)
    Xb, Ub = O.make_tape(rng, D, 1000)
    Sb, _, _ = O.run_mala_sequential_taped(tgt, np.zeros(D), eps, Xb, Ub)
    x0 = Sb[:, -1].astype(np.float32)
    Xi = rng.standard_normal((D, T)).astype(np.float32)
    U = rng.uniform(0.05, 0.95, T)
    x = x0.astype(np.float64)
    for t in range(T):
        yv, _ = O.mala_proposal(tgt, x, eps, Xi[:, t].astype(np.float64))
        la = O.mala_logalpha(tgt, x, yv, eps)
        if t % 12 == 0 and la < -1e-4:
            U[t] = np.exp(la - (4e-6 if t % 24 == 0 else -4e-6))
        if np.log(U[t]) < la:
            x = yv
    return tgt, Chain(DensityModel(dim=D, target=LogisticTarget(Xd, y)), eps, x0, Xi, U.astype(np.float32))


def test_batched_fp64_decisions_near_threshold(monkeypatch):
    """The batched solve re-decides near-threshold columns in FP64 like the
    single-chain solve: every returned decision equals the FP64 decision at
    the returned trajectory (columns within 2e-6 of the threshold there
    excepted, see the single-chain test).  FP32-only batches get some wrong."""
    from ParallelMCMC.Batched import BatchPlan

    pairs = [_near_threshold_chain(s) for s in (81, 82, 83)]

    def run(mode):
        if mode is None:
            monkeypatch.delenv("PMCMC_REFINE", raising=False)
        else:
            monkeypatch.setenv("PMCMC_REFINE", mode)
        plan = BatchPlan([c for _, c in pairs], window=256, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-6,
                         maxiter=600)
        bad = close = 0
        for (tgt, ch), (S, acc, info) in zip(pairs, plan.solve()):
            assert info.converged, f"PMCMC_REFINE={mode}: iters {info.iters}, metric {info.metric}"
            S = S.cpu().numpy().astype(np.float64)
            T = S.shape[1]
            la = O.mala_step_batched_fwd_and_jvp(tgt, O._xbar(ch.x0.astype(np.float64), S), ch.eps,
                                                 ch.Xi.astype(np.float64), np.full(T, 0.5),
                                                 np.ones_like(S)).logalpha
            lu = np.log(ch.U.astype(np.float64))
            margin = np.abs(la - lu)
            got = acc.cpu().numpy().astype(bool)
            checked = margin > 2e-6
            bad += int(np.sum(got[checked] != (lu < la)[checked]))
            close += int(np.sum(checked & (margin < 2e-5)))
        return bad, close

    bad, close = run(None)
    bad_raw, close_raw = run("0")
    print(f"near-threshold columns {close} / {close_raw}; mismatches batched FP64 {bad}, FP32 only {bad_raw}",
          flush=True)
    assert close >= 60
    assert bad == 0
    assert bad_raw > 0
