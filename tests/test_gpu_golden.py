"""GPU DEER solves against the committed sequential-MALA fixtures
(tests/golden/mala_seq_*.npz: FP64 sequential taped MALA from the oracle's C
restatement, cross-checked against the numpy one).  No oracle code runs here:
the FP64 solve must reproduce every decision and the trajectory within 1e-10
(test/test-Deer-vs-MALA.jl:71-108), the FP32 solve (from the FP32-rounded
tape, so against the fixture up to that rounding) every decision and a trace
error within 1e-4, for the diagonal Gaussian, dense Gaussian and logistic
targets."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(kind):
    with np.load(os.path.join(GOLDEN, f"mala_seq_{kind}.npz")) as f:
        return {k: f[k] for k in f.files}


def model(kind, f):
    from ParallelMCMC import DensityModel, DiagGaussianTarget, GaussianTarget, LogisticTarget

    if kind == "diag":
        tgt = DiagGaussianTarget(f["prec"], f["mean"])
    elif kind == "gauss":
        tgt = GaussianTarget(f["Lam"], f["mean"])
    else:
        tgt = LogisticTarget(f["X"], f["y"], float(f["prior_scale"]))
    return DensityModel(dim=len(f["x0"]), target=tgt)


@pytest.mark.parametrize("kind", ["diag", "gauss", "logreg"])
def test_fp64_solve_reproduces_fixture(kind):
    from ParallelMCMC import DEER

    f = load(kind)
    T = f["Xi"].shape[1]
    rec = DEER.TapedRecursion(model(kind, f), float(f["eps"]), f["Xi"], f["U"], precision="f64")
    S, info = DEER.solve(rec, f["x0"], tol_abs=1e-12, tol_rel=1e-11, maxiter=T + 1, damping=1.0, seed=7,
                         return_info=True)
    assert info.converged
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), f["acc"])
    assert np.max(np.abs(S.cpu().numpy() - f["S"])) <= 1e-10 * max(1.0, np.max(np.abs(f["S"])))


@pytest.mark.parametrize("kind", ["diag", "gauss", "logreg"])
def test_fp32_solve_matches_fixture(kind):
    from ParallelMCMC import DEER

    f = load(kind)
    T = f["Xi"].shape[1]
    Xi32 = f["Xi"].astype(np.float32)
    U32 = f["U"].astype(np.float32)
    rec = DEER.TapedRecursion(model(kind, f), float(f["eps"]), Xi32, U32, precision="f32")
    S, info = DEER.solve(rec, f["x0"].astype(np.float32), tol_abs=1e-6, tol_rel=1e-6, maxiter=T + 1, damping=1.0,
                         seed=7, return_info=True)
    assert info.converged
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), f["acc"])
    err = np.max(np.abs(S.cpu().numpy().astype(np.float64) - f["S"]))
    assert err <= 1e-4 * max(1.0, np.max(np.abs(f["S"]))), err
