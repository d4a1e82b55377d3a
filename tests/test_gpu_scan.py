"""GPU parity of the single-pass diagonal affine scan against the oracle's
sequential recurrence (test/test-DEERScan.jl of the reference, plus sizes
that exercise long look-back chains)."""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = ((1, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 17), (8, 31), (8, 64))


def _problem(rng, D, T, dtype=np.float64):
    A = 0.2 * (rng.random((D, T)) - 0.5)
    B = rng.standard_normal((D, T))
    s0 = rng.standard_normal(D)
    return A.astype(dtype), B.astype(dtype), s0.astype(dtype)


def _seq(A, B, s0):
    # vectorised over D, sequential over T, in float64
    A, B = A.astype(np.float64), B.astype(np.float64)
    S = np.empty_like(A)
    s = s0.astype(np.float64)
    for t in range(A.shape[1]):
        s = A[:, t] * s + B[:, t]
        S[:, t] = s
    return S


def _gpu_scan(A, B, s0, torch_dtype):
    import torch
    from ParallelMCMC import DEERScan

    from ParallelMCMC import _device

    dev = torch.device("cuda", 0)
    A_ = _device.to_colmajor(A, torch_dtype, dev)
    B_ = _device.to_colmajor(B, torch_dtype, dev)
    s0_ = _device.to_vec(s0, torch_dtype, dev)
    S = _device.colmajor_empty(A.shape[0], A.shape[1], torch_dtype, dev)
    DEERScan.solve_affine_scan_diag_(S, A_, B_, s0_)
    return S.cpu().numpy()


@pytest.mark.parametrize("D,T", SHAPES)
def test_scan_f64_matches_sequential(D, T):
    import torch

    rng = np.random.default_rng(11 + D * 100 + T)
    A, B, s0 = _problem(rng, D, T)
    S = _gpu_scan(A, B, s0, torch.float64)
    assert np.allclose(S, _seq(A, B, s0), rtol=1e-12, atol=1e-12)
    assert np.allclose(S, O.solve_affine_seq(A, B, s0), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("D,T", [(1, 1), (1, 17), (4, 16), (8, 31), (16, 64), (128, 1024)])
def test_scan_f32_matches_sequential(D, T):
    import torch

    rng = np.random.default_rng(1 + D + T)
    A, B, s0 = _problem(rng, D, T, np.float32)
    S = _gpu_scan(A, B, s0, torch.float32)
    assert np.allclose(S, _seq(A, B, s0), rtol=1e-5, atol=1e-5)


def test_scan_workspace_reuse_and_nan_fill():
    import torch
    from ParallelMCMC import DEERScan, _device

    rng = np.random.default_rng(3)
    A, B, s0 = _problem(rng, 8, 64)
    dev = torch.device("cuda", 0)
    A_, B_ = (_device.to_colmajor(M, torch.float64, dev) for M in (A, B))
    S = _device.colmajor_empty(8, 64, torch.float64, dev)
    ws = DEERScan.AffineScanWorkspace(A_)
    out = DEERScan.solve_affine_scan_diag_(S, A_, B_, s0, ws)
    assert out is S
    S.fill_(float("nan"))
    DEERScan.solve_affine_scan_diag_(S, A_, B_, s0, ws)
    assert np.allclose(S.cpu().numpy(), _seq(A, B, s0), rtol=1e-12, atol=1e-12)


def test_scan_edge_cases():
    import torch

    rng = np.random.default_rng(4)
    A, B, s0 = _problem(rng, 5, 40)
    # A = 0 => S = B ; B = 0 => running products times s0 ; T = 1
    assert np.array_equal(_gpu_scan(np.zeros_like(A), B, s0, torch.float64), B)
    Ap = 0.5 + 0.5 * rng.random((5, 40))
    assert np.allclose(_gpu_scan(Ap, np.zeros_like(B), s0, torch.float64),
                       np.cumprod(Ap, axis=1) * s0[:, None], rtol=1e-12)
    assert np.allclose(_gpu_scan(A[:, :1], B[:, :1], s0, torch.float64), A[:, :1] * s0[:, None] + B[:, :1])


def test_scan_shape_and_alias_errors():
    import torch
    from ParallelMCMC import ArgumentError, DEERScan, DimensionMismatch, _device

    dev = torch.device("cuda", 0)
    A = _device.colmajor_empty(3, 5, torch.float64, dev).normal_()
    B = _device.colmajor_empty(3, 5, torch.float64, dev).normal_()
    S = _device.colmajor_empty(3, 5, torch.float64, dev)
    with pytest.raises(DimensionMismatch):
        DEERScan.solve_affine_scan_diag_(S, A, B[:, :4], torch.zeros(3, dtype=torch.float64, device=dev))
    with pytest.raises(DimensionMismatch):
        DEERScan.solve_affine_scan_diag_(S, A, B, torch.zeros(4, dtype=torch.float64, device=dev))
    with pytest.raises(ArgumentError):
        DEERScan.solve_affine_scan_diag_(A, A, B, torch.zeros(3, dtype=torch.float64, device=dev))
    # T == 0 is a no-op
    Z = _device.colmajor_empty(3, 0, torch.float64, dev)
    assert DEERScan.solve_affine_scan_diag_(Z, Z, Z, torch.zeros(3, dtype=torch.float64, device=dev)) is Z


@pytest.mark.parametrize("D,T", [(1, 1 << 20), (3, 100003), (100, 4096), (256, 65536), (1000, 300)])
def test_scan_large_sizes(D, T):
    """Many tiles -> long look-back chains.  Multipliers near 1 so that the
    state carries information across the whole sequence."""
    import torch

    rng = np.random.default_rng(D + T)
    A = (0.999 + 0.001 * rng.random((D, T))).astype(np.float32)
    B = (0.01 * rng.standard_normal((D, T))).astype(np.float32)
    s0 = rng.standard_normal(D).astype(np.float32)
    S = _gpu_scan(A, B, s0, torch.float32)
    ref = _seq(A, B, s0)
    err = np.max(np.abs(S - ref)) / max(1.0, np.max(np.abs(ref)))
    assert err < 2e-4, err  # float32 recurrence over up to 1e6 steps
    # float64 on a subset of sizes: near-exact
    if D * T <= 4_000_000:
        S64 = _gpu_scan(A.astype(np.float64), B.astype(np.float64), s0.astype(np.float64), torch.float64)
        assert np.allclose(S64, ref, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("D,T,dt", [(1, 1 << 20, "f64"), (64, 65536, "f64"), (300, 4000, "f64"),
                                    (64, 1 << 18, "f32"), (7, 123457, "f32")])
def test_scan_long_carry_chained_vs_two_pass(D, T, dt, monkeypatch):
    """A near 1 carries every tile's state into all later tiles, so a wrong
    look-back composition anywhere shows up downstream.  The single-pass
    (decoupled look-back) scan matches the sequential recurrence and the
    two-pass reduce/apply scan."""
    import torch

    rng = np.random.default_rng(D + T)
    dtype = np.float64 if dt == "f64" else np.float32
    A = rng.uniform(0.999, 1.0005, (D, T)).astype(dtype)
    B = (1e-3 * rng.standard_normal((D, T))).astype(dtype)
    s0 = rng.standard_normal(D).astype(dtype)
    tdt = torch.float64 if dt == "f64" else torch.float32
    monkeypatch.setenv("PMCMC_SCAN_CHAINED", "1")
    S1 = _gpu_scan(A, B, s0, tdt)
    monkeypatch.delenv("PMCMC_SCAN_CHAINED", raising=False)
    S2 = _gpu_scan(A, B, s0, tdt)
    ref = _seq(A, B, s0)
    tol = 1e-10 if dt == "f64" else 1e-3
    scale = np.max(np.abs(ref))
    print(f"D={D} T={T} {dt}: chained err {np.max(np.abs(S1 - ref)) / scale:.2e}, "
          f"two-pass err {np.max(np.abs(S2 - ref)) / scale:.2e}", flush=True)
    assert np.max(np.abs(S1 - ref)) <= tol * scale
    assert np.max(np.abs(S2 - ref)) <= tol * scale
