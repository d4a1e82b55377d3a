"""Autotuner on the GPU: every level-0 shape the tuner can pick gives the same
results (diagonal scan vs the sequential recurrence, DEER solves vs the FP64
sequential MALA run), and the tuners' winners land in the table and are used
by later scans and workspaces."""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _clean_table():
    from ParallelMCMC import Tune

    Tune.clear()
    yield
    Tune.clear()


def _scan_case(rng, D, T):
    A = rng.uniform(0.5, 1.0, (D, T))
    B = rng.standard_normal((D, T))
    s0 = rng.standard_normal(D)
    return A, B, s0, O.solve_affine_seq(A, B, s0)


def _reg_shapes(D, fp64=True):
    """Register-tile shapes (tc, parts, form) the library offers at D (16 <= D
    <= 256, split into row blocks of at most 64 rows): tc = 64 or 128 (and 256
    in FP32, one tile per workgroup)."""
    if not 16 <= D <= 256:
        return set()
    out = {(64, 1, 2), (128, 1, 2), (64, 1, 3), (128, 1, 3)}
    if not fp64:
        out.add((256, 1, 2))
    return out


@pytest.mark.parametrize("D,T", [(64, 20000), (5, 30001), (100, 3000), (40, 7001), (256, 4099), (131, 5003)])
def test_every_candidate_scan_shape_matches_sequential(D, T):
    import torch

    from ParallelMCMC import DEERScan, Tune

    rng = np.random.default_rng(D + T)
    A, B, s0, ref = _scan_case(rng, D, T)
    pmax = max(1, 256 // D)
    rule_tc = max(1, 2048 // D)  # FP64 tile rule
    shapes = set()
    for tc in (max(1, rule_tc // 4), max(1, rule_tc // 2), rule_tc, 2 * rule_tc):
        for p in sorted({pmax, max(1, pmax // 2), 1}):
            if p <= tc:
                shapes.add((tc, p, False))
        shapes.add((tc, min(pmax, tc), 1))
    reg = _reg_shapes(D)  # register-resident tiles, one per workgroup or persistent
    shapes.update(reg)
    checked = 0
    reg_checked = 0
    for tc, p, sp in sorted(shapes):
        try:
            Tune.set_entry("scan_diag", torch.float64, D, T, tc=tc, parts=p, single_pass=sp)
        except Exception:
            continue  # does not fit the LDS budget: the tuner never proposes it
        S = DEERScan.solve_affine_scan_diag(A, B, s0).cpu().numpy()
        assert np.max(np.abs(S - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), (tc, p, sp)
        checked += 1
        reg_checked += (tc, p, sp) in reg
    assert checked >= 6
    assert reg_checked == len(reg)  # every register shape is accepted and checked


def test_scan_tune_records_and_applies_its_winner():
    import torch

    from ParallelMCMC import DEERScan, Tune

    D, T = 64, 1 << 17
    best = Tune.tune_scan(D, T, torch.float32, reps=3)
    assert best["tc"] >= 1 and best["parts"] >= 1 and best["us"] > 0
    assert Tune.lookup("scan_diag", torch.float32, D, T) == best
    rng = np.random.default_rng(4)
    A, B, s0, ref = _scan_case(rng, D, T)
    S = DEERScan.solve_affine_scan_diag(A.astype(np.float32), B.astype(np.float32), s0.astype(np.float32))
    ref32 = O.solve_affine_seq(A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64),
                               s0.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(S.cpu().numpy() - ref32)) <= 1e-5 * np.max(np.abs(ref32))


def _logistic(seed, N, D, T, eps, prec):
    from ParallelMCMC import DEER, DensityModel, LogisticTarget

    rng = np.random.default_rng(seed)
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    tgt = O.LogReg(Xd, y)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y))
    dt = np.float64 if prec == "f64" else np.float32
    x0 = (0.1 * rng.standard_normal(D)).astype(dt)
    Xi, U = O.make_tape(rng, D, T, dt)
    ref, acc, _ = O.run_mala_sequential_taped(tgt, x0.astype(np.float64), eps, Xi.astype(np.float64),
                                              U.astype(np.float64))
    return DEER.TapedRecursion(model, eps, Xi, U, precision=prec), x0, ref, acc


@pytest.mark.parametrize("shape", [(8, 1), (8, 2), (16, 2), (4, 4), (32, 1)])
def test_deer_solve_under_forced_tile_shapes(shape):
    """FP64 logistic solve (D=20, T=700) with each level-0 shape: trajectory
    within 1e-10 of sequential MALA and identical decisions."""
    import torch

    from ParallelMCMC import DEER, Tune

    tc, p = shape
    rec, x0, ref, acc = _logistic(5, 400, 20, 700, 1e-2, "f64")
    Tune.set_entry("deer", torch.float64, 20, 700, tc=tc, parts=p)
    S, info = DEER.solve(rec, x0, tol_abs=1e-12, tol_rel=1e-11, maxiter=700, damping=1.0, seed=3, return_info=True)
    assert info.converged
    assert np.max(np.abs(S.cpu().numpy() - ref)) < 1e-10
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc)


def test_deer_tune_then_solve():
    """Tune the FP32 logistic workload's tile shape, then solve with it."""
    from ParallelMCMC import DEER, Tune

    rec, x0, ref, acc = _logistic(7, 2000, 100, 512, 5e-3, "f32")
    ws = DEER.DEERWorkspace(100, 512, rec.precision)
    best = Tune.tune_solve(rec, x0, iters=4, workspace=ws, seed=11)
    assert best["tc"] >= 1 and best["us"] > 0
    assert Tune.lookup("deer", "f32", 100, 512)["tc"] == best["tc"]
    S, info = DEER.solve(rec, x0, tol_abs=1e-6, tol_rel=1e-6, maxiter=500, damping=1.0, seed=11, return_info=True,
                         workspace=ws, copy_result=True)
    assert info.converged
    assert O.trace_error(S.cpu().numpy().astype(np.float64), ref) < 1e-5
    assert np.array_equal(info.accepted.cpu().numpy().astype(bool), acc)


@pytest.mark.parametrize("tc,form", [(64, 2), (128, 2), (256, 2), (64, 3), (128, 3)])
def test_register_tile_scan_long_carry_fp32(tc, form):
    """Register-tile single-pass scan (D = 64, FP32; one tile per workgroup, or
    persistent workgroups) on a ragged T with A near 1, so each tile's incoming
    state depends on many predecessors' maps."""
    import torch

    from ParallelMCMC import DEERScan, Tune

    D, T = 64, 300017
    rng = np.random.default_rng(tc)
    A = rng.uniform(0.999, 1.0, (D, T)).astype(np.float32)
    B = rng.standard_normal((D, T)).astype(np.float32)
    s0 = rng.standard_normal(D).astype(np.float32)
    ref = np.empty((D, T))
    x = s0.astype(np.float64)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    for t in range(T):
        x = A64[:, t] * x + B64[:, t]
        ref[:, t] = x
    Tune.set_entry("scan_diag", torch.float32, D, T, tc=tc, parts=1, single_pass=form)
    S = DEERScan.solve_affine_scan_diag(A, B, s0).cpu().numpy().astype(np.float64)
    assert np.max(np.abs(S - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.parametrize("D,tc,T,form", [(64, 128, 5000, 2), (64, 256, 64 * 256 * 3 + 17, 2), (64, 64, 1_200_000, 2),
                                         (64, 128, 5000, 3), (64, 64, 1_200_000, 3), (64, 128, 64 * 128 * 5 + 3, 3),
                                         (72, 64, 1_100_000, 2), (128, 64, 64 * 64 * 5 + 3, 3),
                                         (200, 128, 300_001, 3), (256, 256, 20_000, 2), (48, 64, 9_999, 3)])
def test_register_tile_scan_exact_running_sum(D, tc, T, form):
    """Register-tile scans (form 2: one tile per workgroup; 3: persistent
    workgroups) with A = 1 and integer B (|S| < 2^24, so FP32 is exact): S
    must equal s0 + cumsum(B) exactly.  Covers one partial group of
    tiles (T = 5000, 40 tiles), whole groups, and more than 256 groups of 64
    tiles (T = 1.2e6 at 64 columns per tile: 293 groups, a multi-window
    look-back); D above 64 runs as row blocks (D = 72: two blocks of 36 rows,
    more than 256 groups each; 128, 200, 256: two to four blocks), D = 48 has
    fewer rows than lanes."""
    import torch

    from ParallelMCMC import DEERScan, Tune

    rng = np.random.default_rng(T)
    A = np.ones((D, T), dtype=np.float32)
    B = rng.integers(-1, 2, (D, T)).astype(np.float32)
    s0 = rng.integers(-50, 50, D).astype(np.float32)
    ref = s0[:, None].astype(np.float64) + np.cumsum(B.astype(np.float64), axis=1)
    Tune.set_entry("scan_diag", torch.float32, D, T, tc=tc, parts=1, single_pass=form)
    S = DEERScan.solve_affine_scan_diag(A, B, s0).cpu().numpy().astype(np.float64)
    assert np.array_equal(S, ref)
