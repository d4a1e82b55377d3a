"""GPU tests of the block trajectory and log-densities taken from the final
Newton iteration (DEER.last_trajectory, pmcmc_deer_last_trajectory) on the
GEMM targets, and of the sampler that uses them (the reference evaluates
_trajectory_logps separately, src/interface.jl:517-524, 586-594).

* FP64: the assembled trajectory FT(S_k) is sequential MALA within 1e-10
  and its log-densities equal the oracle's log p of those states to 1e-10
  relative; it is within the solve's tolerance of the returned iterate;
* FP32: log-densities within 1e-5 relative of the oracle's on the same
  states;
* sample() on a logistic model: every block is sequential MALA on its tape
  and every kept sample's logp is log p of that sample;
* a solve on a diagonal target leaves nothing to assemble: ArgumentError.
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


def _gauss(seed, D, cond=2.0):
    from ParallelMCMC import DensityModel, GaussianTarget

    rng = np.random.default_rng(seed)
    Lam, ev = O.make_corr_gaussian(rng, D, cond)
    mu = 0.3 * rng.standard_normal(D)
    return rng, O.Gaussian(Lam, mu), DensityModel(dim=D, target=GaussianTarget(Lam, mu)), ev


@pytest.mark.parametrize("kind,jac", [("logistic", "stoch_diag"), ("logistic", "diag"), ("gauss", "stoch_diag"),
                                      ("gauss", "full")])
def test_last_trajectory_f64(kind, jac):
    from ParallelMCMC import DEER

    if kind == "logistic":
        rng, tgt, model = _logistic(41, 300, 9, 0.8)
        D, T, eps = 9, 200, 0.02
        x0 = 0.3 * rng.standard_normal(D)
    else:
        rng, tgt, model, ev = _gauss(42, 12)
        D, T = 12, 160
        eps = 0.9 / ev.max()
        x0 = tgt.mean + rng.standard_normal(D)
    Xi, U = O.make_tape(rng, D, T)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    rec = DEER.TapedRecursion(model, eps, Xi, U)
    ws = DEER.DEERWorkspace(D, T)
    S, info = DEER.solve(rec, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=400, jacobian=jac, damping=1.0,
                         return_info=True, workspace=ws)
    assert info.converged
    S_it = S.cpu().numpy()
    S_ft, lp = DEER.last_trajectory(ws)
    S_ft, lp = S_ft.cpu().numpy(), lp.cpu().numpy()
    assert np.max(np.abs(S_ft - ref)) < 1e-10
    assert np.max(np.abs(S_ft - S_it)) <= 10 * (1e-12 + 1e-11 * np.max(np.abs(S_it)))
    lp_ref = tgt.logp_batch(S_ft)
    assert np.max(np.abs(lp - lp_ref)) <= 1e-10 * max(1.0, np.max(np.abs(lp_ref)))


@pytest.mark.parametrize("prec", ["f32", "bf16_jac"])
def test_last_trajectory_f32_logistic(prec):
    from ParallelMCMC import DEER

    rng, tgt, model = _logistic(7, 2000, 100)
    D, T, eps = 100, 512, 5e-3
    x0 = (0.1 * rng.standard_normal(D)).astype(np.float32)
    Xi, U = O.make_tape(rng, D, T, np.float32)
    ref, acc_ref, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                                  U.astype(np.float64))
    rec = DEER.TapedRecursion(model, eps, Xi, U, precision=prec)
    ws = DEER.DEERWorkspace(D, T, rec.precision)
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=500, return_info=True, workspace=ws)
    assert info.converged
    S_ft, lp = DEER.last_trajectory(ws)
    S_ft = S_ft.cpu().numpy().astype(np.float64)
    lp = lp.cpu().numpy()
    assert O.trace_error(S_ft, ref) < 1e-5
    lp_ref = tgt.logp_batch(S_ft)
    rel = np.max(np.abs(lp - lp_ref)) / np.max(np.abs(lp_ref))
    print(f"{prec}: logp relative error {rel:.2e}", flush=True)
    assert rel < 1e-5


def test_sampler_blocks_logistic_logps_from_last_iteration():
    from ParallelMCMC import ParallelMALASampler, sample
    from ParallelMCMC.interface import make_mala_tape_block

    rng0, tgt, model = _logistic(5, 250, 6)
    D, N, T, eps = 6, 150, 64, 0.03
    sampler = ParallelMALASampler(eps, T=T, tol_abs=1e-13, tol_rel=1e-12)
    samples, infos = sample(np.random.default_rng(77), model, sampler, N, return_infos=True)
    assert len(samples) == N and all(i.converged for i in infos)
    got = samples.array()
    rng = np.random.default_rng(77)
    x = rng.standard_normal(D)
    row = 0
    while row < N:
        Xi, U = make_mala_tape_block(rng, D, T, sampler.precision)
        rng.integers(0, 2**63 - 1)
        ref, _, _ = O.run_mala_sequential_taped(tgt, x, eps, Xi, U)
        nkeep = min(T, N - row)
        assert np.max(np.abs(got[row:row + nkeep].T - ref[:, :nkeep])) < 1e-10
        x = got[row + nkeep - 1].copy()
        row += nkeep
    lps = np.array([samples[i].logp for i in range(N)])
    assert np.max(np.abs(lps - tgt.logp_batch(got.T))) < 1e-9


def test_last_trajectory_needs_a_gemm_solve():
    from ParallelMCMC import DEER, DensityModel, StdNormalTarget
    from ParallelMCMC._lib import ArgumentError

    D, T = 3, 16
    rng = np.random.default_rng(0)
    Xi, U = O.make_tape(rng, D, T)
    rec = DEER.TapedRecursion(DensityModel(dim=D, target=StdNormalTarget(D)), 0.1, Xi, U)
    ws = DEER.DEERWorkspace(D, T)
    DEER.solve(rec, np.zeros(D), workspace=ws)
    with pytest.raises(ArgumentError):
        DEER.last_trajectory(ws)
