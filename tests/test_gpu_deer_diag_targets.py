"""GPU parity of the fused DEER iteration for diagonal-Hessian targets
against the oracle: the batched MALA step + JVP, one Newton update, the full
solve (same iteration count, same accept decisions, trajectory equal to the
sequential taped MALA run) and the block sampler (test/test-Deer-vs-MALA.jl,
test/test-GPU-MALA.jl, test/test-GPU-DEER.jl of the reference)."""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def _model(tgt_np):
    from ParallelMCMC import DensityModel, DiagGaussianTarget, StdNormalTarget

    if np.all(tgt_np.prec == 1) and np.all(tgt_np.mean == 0):
        return DensityModel(dim=tgt_np.dim, target=StdNormalTarget(tgt_np.dim))
    return DensityModel(dim=tgt_np.dim, target=DiagGaussianTarget(tgt_np.prec, tgt_np.mean))


def _targets(rng, D):
    return [O.DiagGaussian.std_normal(D), O.DiagGaussian(rng.uniform(0.5, 3.0, D), rng.standard_normal(D))]


@pytest.mark.parametrize("D,T", [(1, 7), (3, 16), (7, 33), (100, 64), (300, 5)])
def test_fwd_and_jvp_f64(D, T):
    from ParallelMCMC import MALA

    rng = np.random.default_rng(D * 7 + T)
    for tgt in _targets(rng, D):
        X, Xi = rng.standard_normal((2, D, T))
        Z = np.where(rng.random((D, T)) < 0.5, -1.0, 1.0)
        U = rng.uniform(0.01, 0.99, T)
        ref = O.mala_step_batched_fwd_and_jvp(tgt, X, 0.3, Xi, U, Z)
        FT, Jt, acc, la = MALA.mala_step_batched_fwd_and_jvp(_model(tgt), X, 0.3, Xi, U, Z, return_logalpha=True)
        assert np.array_equal(acc.cpu().numpy().astype(bool), ref.acc)
        assert np.allclose(la.cpu().numpy(), ref.logalpha, rtol=1e-12, atol=1e-12)
        assert np.allclose(FT.cpu().numpy(), ref.FT, rtol=1e-13, atol=1e-13)
        assert np.allclose(Jt.cpu().numpy(), ref.Jt, rtol=1e-11, atol=1e-11)


def test_fwd_and_jvp_f32_identical_decisions():
    # test/test-GPU-MALA.jl:74-95: identical accept decisions, states within 1e-5
    from ParallelMCMC import MALA

    rng = np.random.default_rng(42)
    D, T = 3, 4096
    tgt = O.DiagGaussian.std_normal(D)
    X, Xi = (rng.standard_normal((D, T)).astype(np.float32) for _ in range(2))
    U = rng.random(T).astype(np.float32)
    Z = np.where(rng.random((D, T)) < 0.5, -1.0, 1.0).astype(np.float32)
    ref = O.mala_step_batched_fwd_and_jvp(tgt, X.astype(np.float64), 0.1, Xi.astype(np.float64), U.astype(np.float64), Z.astype(np.float64))
    FT, Jt, acc = MALA.mala_step_batched_fwd_and_jvp(_model(tgt), X, 0.1, Xi, U, Z, precision="f32")
    assert np.array_equal(acc.cpu().numpy().astype(bool), ref.acc)
    assert np.allclose(FT.cpu().numpy(), ref.FT, rtol=1e-5, atol=1e-6)
    assert np.allclose(Jt.cpu().numpy(), ref.Jt, rtol=1e-4, atol=1e-5)


def test_forced_accept_reject_batched():
    # test/test-GPU-MALA.jl:122-158: u ~ 1 rejects leave X unchanged, u ~ 0 accepts all
    from ParallelMCMC import MALA

    rng = np.random.default_rng(5)
    D, T = 4, 64
    tgt = O.DiagGaussian.std_normal(D)
    X, Xi, Z = rng.standard_normal((3, D, T))
    FT, _, acc = MALA.mala_step_batched_fwd_and_jvp(_model(tgt), X, 0.1, Xi, np.full(T, 1 - 1e-6), Z)
    acc = acc.cpu().numpy().astype(bool)
    assert np.array_equal(FT.cpu().numpy()[:, ~acc], X[:, ~acc]) and acc.sum() < T
    _, _, acc0 = MALA.mala_step_batched_fwd_and_jvp(_model(tgt), X, 0.001, Xi, np.full(T, 1e-7), Z)
    assert acc0.cpu().numpy().all()
    _, _, accz = MALA.mala_step_batched_fwd_and_jvp(_model(tgt), X, 0.5, Xi, np.zeros(T), Z)
    assert accz.cpu().numpy().all()  # u = 0: log u = -Inf accepts (batched path)


@pytest.mark.parametrize("jac,probes", [("stoch_diag", 1), ("stoch_diag", 3), ("diag", 1)])
def test_single_newton_update_matches_oracle(jac, probes):
    from ParallelMCMC import DEER

    rng = np.random.default_rng(17)
    D, T, eps = 5, 50, 0.2
    for tgt in _targets(rng, D):
        s0 = rng.standard_normal(D)
        Xi, U = O.make_tape(rng, D, T)
        S_in = rng.standard_normal((D, T))
        ref, acc, _ = O.deer_update(tgt, eps, Xi, U, s0, S_in, jac, 0.7, probes, seed=99, it=0)
        rec = DEER.TapedRecursion(_model(tgt), eps, Xi, U)
        S = DEER.deer_update(rec, s0, S_in, jacobian=jac, damping=0.7, probes=probes, seed=99)
        assert np.allclose(S.cpu().numpy(), ref, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("jac,probes,eps,seed", [("diag", 1, 0.05, 12345), ("stoch_diag", 2, 0.15, 777),
                                                  ("stoch_diag", 1, 0.35, 3)])
def test_deer_solve_recovers_sequential_mala_f64(jac, probes, eps, seed):
    # test/test-Deer-vs-MALA.jl:71-108, 161-215 at the 1e-10 FP64 bar
    from ParallelMCMC import DEER

    rng = np.random.default_rng(seed)
    D, T = 3, 32
    tgt = O.DiagGaussian.std_normal(D)
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    S_o, info_o = O.deer_solve(tgt, eps, Xi, U, x0, tol_abs=1e-13, tol_rel=1e-12, maxiter=400,
                               jacobian=jac, damping=0.5, probes=probes, seed=42)
    rec = DEER.TapedRecursion(_model(tgt), eps, Xi, U)
    S, info = DEER.solve(rec, x0, tol_abs=1e-13, tol_rel=1e-12, maxiter=400, jacobian=jac, damping=0.5,
                         probes=probes, seed=42, return_info=True)
    S = S.cpu().numpy()
    assert info.converged and info_o.converged
    assert info.iters <= info_o.iters
    assert np.max(np.abs(S - ref)) < 1e-10
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)


def test_deer_solve_large_diag_gaussian_f64_and_f32():
    """D=64, T=4096, anisotropic diagonal Gaussian: many tiles per iteration."""
    from ParallelMCMC import DEER

    rng = np.random.default_rng(8)
    D, T, eps = 64, 4096, 0.05
    tgt = O.DiagGaussian(rng.uniform(0.5, 4.0, D), 0.3 * rng.standard_normal(D))
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    rec = DEER.TapedRecursion(_model(tgt), eps, Xi, U)
    S, info = DEER.solve(rec, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=1000, damping=1.0, seed=1,
                         return_info=True)
    assert info.converged
    assert np.max(np.abs(S.cpu().numpy() - ref)) < 1e-10
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc_ref)
    # float32 tape (rounded) vs the FP64 oracle on the same rounded tape
    Xi32, U32 = Xi.astype(np.float32), U.astype(np.float32)
    ref32, acc32, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float32).astype(np.float64), eps,
                                                  Xi32.astype(np.float64), U32.astype(np.float64))
    rec32 = DEER.TapedRecursion(_model(tgt), eps, Xi32, U32, precision="f32")
    S32, info32 = DEER.solve(rec32, x0.astype(np.float32), tol_abs=1e-6, tol_rel=1e-6, maxiter=1000, damping=1.0,
                             seed=1, return_info=True)
    assert info32.converged
    err = np.max(np.abs(S32.cpu().numpy() - ref32)) / np.max(np.abs(ref32))
    assert err < 1e-5, err
    assert np.array_equal(info32.accepted.cpu().numpy().astype(bool), acc32)


def test_nonconvergence_reported():
    from ParallelMCMC import DEER

    rng = np.random.default_rng(9)
    D, T = 4, 200
    tgt = O.DiagGaussian.std_normal(D)
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    rec = DEER.TapedRecursion(_model(tgt), 0.3, Xi, U)
    S, info = DEER.solve(rec, x0, tol_abs=1e-12, tol_rel=1e-12, maxiter=1, damping=0.5, return_info=True)
    assert not info.converged and info.iters == 1
    _, io = O.deer_solve(tgt, 0.3, Xi, U, x0, tol_abs=1e-12, tol_rel=1e-12, maxiter=1, damping=0.5)
    assert abs(info.metric - io.metric) <= 1e-9 * io.metric


def test_nan_never_converges():
    """A NaN entering the Newton update (here through the JVP of a NaN
    proposal, which the accept test itself rejects) fails the convergence
    test, as maximum() makes it do in src/DEER/DEER.jl:567-571."""
    from ParallelMCMC import DEER

    rng = np.random.default_rng(10)
    D, T = 3, 100
    tgt = O.DiagGaussian.std_normal(D)
    x0 = rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    Xi[1, 40] = np.nan
    _, io = O.deer_solve(tgt, 0.2, Xi, U, x0, tol_abs=1e-6, tol_rel=1e-5, maxiter=5, damping=1.0)
    assert not io.converged and np.isnan(io.metric)
    rec = DEER.TapedRecursion(_model(tgt), 0.2, Xi, U)
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-5, maxiter=5, damping=1.0, return_info=True)
    assert not info.converged and info.iters == 5 and np.isnan(info.metric)


def test_c0_sampler_blocks_match_sequential():
    """C0: 2-D standard normal, eps=0.1, T=64, :stoch_diag, FP64, 500 samples
    via sample(); every block equals sequential taped MALA on its tape."""
    from ParallelMCMC import DensityModel, ParallelMALASampler, StdNormalTarget, sample
    from ParallelMCMC.interface import make_mala_tape_block

    D, N = 2, 500
    model = DensityModel(dim=D, target=StdNormalTarget(D))
    sampler = ParallelMALASampler(0.1, T=64, tol_abs=1e-13, tol_rel=1e-12)
    samples, infos = sample(np.random.default_rng(2024), model, sampler, N, return_infos=True)
    assert len(samples) == N and all(i.converged for i in infos)
    rng = np.random.default_rng(2024)
    x = rng.standard_normal(D)
    tgt = O.DiagGaussian.std_normal(D)
    got = samples.array()
    row = 0
    while row < N:
        Xi, U = make_mala_tape_block(rng, D, 64, sampler.precision)
        rng.integers(0, 2**63 - 1)
        ref, _, _ = O.run_mala_sequential_taped(tgt, x, 0.1, Xi, U)
        nkeep = min(64, N - row)
        assert np.max(np.abs(got[row:row + nkeep].T - ref[:, :nkeep])) < 1e-10
        x = got[row + nkeep - 1].copy()
        row += nkeep
    # logps of the kept samples
    assert abs(samples[3].logp - (-0.5 * float(samples[3].x @ samples[3].x))) < 1e-12


def test_step_interface_walks_and_resolves():
    from ParallelMCMC import DensityModel, ParallelMALASampler, StdNormalTarget, step

    model = DensityModel(dim=3, target=StdNormalTarget(3))
    sampler = ParallelMALASampler(0.2, T=8)
    rng = np.random.default_rng(1)
    tr, st = step(rng, model, sampler, initial_params=np.zeros(3))
    xs = [tr.x]
    for _ in range(20):
        tr, st = step(rng, model, sampler, st)
        xs.append(tr.x)
    assert st.t == 20 % 8
    assert all(np.all(np.isfinite(x)) for x in xs)
