"""The owned matmul and reductions (src/ParallelMCMC.jl:23-25) on the MFMA
GEMM and the fixed-order reductions, against numpy; their derivative rules
(ext/EnzymeExt.jl:78-284) through torch autograd: reverse pullbacks by
gradcheck (FP64 finite differences) and forward JVPs against the analytic
dA B + A dB.  Mirrors test/test-Owned-Matmul.jl of the reference."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(a):
    import torch

    return torch.tensor(a, device="cuda")


def _bound32(A, B):
    # FP32 products of k terms: |error| <= ~k eps32 (|A| |B|)
    k = A.shape[1]
    return 4 * k * 2.0 ** -24 * (np.abs(A) @ np.abs(B)) + 1e-30


@pytest.mark.parametrize("m,k,n", [(4, 3, 5), (1, 1, 1), (64, 16, 64), (67, 130, 45), (300, 257, 513), (5, 0, 7)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_matmul_values(m, k, n, dt):
    from ParallelMCMC import pmcmc_matmul

    rng = np.random.default_rng(m * 1000 + k * 10 + n)
    A = rng.standard_normal((m, k)).astype(dt)
    B = rng.standard_normal((k, n)).astype(dt)
    got = pmcmc_matmul(_t(A), _t(B)).cpu().numpy()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert got.shape == (m, n)
    if dt == np.float64:
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    else:
        assert np.all(np.abs(got - ref) <= _bound32(A.astype(np.float64), B.astype(np.float64)))


def test_matmul_vector_forms_match_julia():
    """A * b (matrix-vector) and a * B with B one row (outer product), the
    reference's first test case (4x3 times 3x5 and times a 3-vector)."""
    from ParallelMCMC import ArgumentError, pmcmc_matmul

    rng = np.random.default_rng(0)
    A = rng.standard_normal((4, 3))
    B = rng.standard_normal((3, 5))
    b = rng.standard_normal(3)
    assert np.allclose(pmcmc_matmul(_t(A), _t(B)).cpu().numpy(), A @ B, rtol=1e-12)
    yb = pmcmc_matmul(_t(A), _t(b)).cpu().numpy()
    assert yb.shape == (4,) and np.allclose(yb, A @ b, rtol=1e-12)
    a = rng.standard_normal(4)
    r = rng.standard_normal((1, 6))
    assert np.allclose(pmcmc_matmul(_t(a), _t(r)).cpu().numpy(), np.outer(a, r[0]), rtol=1e-12)
    with pytest.raises(ArgumentError):
        pmcmc_matmul(_t(a), _t(a))


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_matmul_abi_transposes_and_leading_dims(ta, tb, prec):
    """C = op(A) op(B) straight through the C ABI, column-major, with leading
    dimensions larger than the stored rows and a C block inside a larger array."""
    import torch

    from ParallelMCMC import _device, _lib

    dt = torch.float64 if prec == "f64" else torch.float32
    code = _lib.PREC_F64 if prec == "f64" else _lib.PREC_F32
    rng = np.random.default_rng(3 + 2 * ta + tb)
    m, n, k = 37, 70, 19
    Aop = rng.standard_normal((m, k))
    Bop = rng.standard_normal((k, n))
    As = Aop.T if ta else Aop  # the stored (column-major) matrices
    Bs = Bop.T if tb else Bop
    lda, ldb, ldc = As.shape[0] + 3, Bs.shape[0] + 5, m + 2
    Abuf = np.zeros((lda, As.shape[1]))
    Abuf[: As.shape[0]] = As
    Bbuf = np.zeros((ldb, Bs.shape[1]))
    Bbuf[: Bs.shape[0]] = Bs
    Cbuf = np.full((ldc, n), 7.0)
    # column-major device buffers: torch (cols, ld) row-major == (ld, cols) column-major
    dA = torch.tensor(Abuf.T.copy(), dtype=dt, device="cuda")
    dB = torch.tensor(Bbuf.T.copy(), dtype=dt, device="cuda")
    dC = torch.tensor(Cbuf.T.copy(), dtype=dt, device="cuda")
    ctx = _device.context(0)
    ctx.call("pmcmc_matmul", code, ta, tb, m, n, k, _device.ptr(dA), lda, _device.ptr(dB), ldb, _device.ptr(dC), ldc)
    Cgot = dC.cpu().numpy().astype(np.float64).T
    ref = Aop.astype(np.float32).astype(np.float64) @ Bop.astype(np.float32).astype(np.float64) if prec == "f32" \
        else Aop @ Bop
    tol = 1e-12 if prec == "f64" else 1e-4
    assert np.allclose(Cgot[:m], ref, rtol=tol, atol=tol)
    assert np.all(Cgot[m:] == 7.0)  # rows beyond m untouched


def test_matmul_abi_rejects_bad_leading_dims():
    from ParallelMCMC import _device, _lib

    ctx = _device.context(0)
    lib = _lib.load()
    rc = lib.pmcmc_matmul(ctx.h, _lib.PREC_F64, 0, 0, 8, 4, 3, C.c_void_p(16), 7, C.c_void_p(16), 3, C.c_void_p(16), 8)
    assert rc == 2  # DimensionMismatch: lda < m
    rc = lib.pmcmc_matmul(ctx.h, _lib.PREC_BF16_JAC, 0, 0, 8, 4, 3, C.c_void_p(16), 8, C.c_void_p(16), 3,
                          C.c_void_p(16), 8)
    assert rc == 1


def test_matmul_reverse_rule_gradcheck():
    """Pullbacks dA = dY B^T and dB = A^T dY (the transposed GEMMs) against
    FP64 central finite differences."""
    import torch

    from ParallelMCMC import pmcmc_matmul

    rng = np.random.default_rng(5)
    A = torch.tensor(rng.standard_normal((5, 4)), device="cuda", requires_grad=True)
    B = torch.tensor(rng.standard_normal((4, 3)), device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(pmcmc_matmul, (A, B), eps=1e-6, atol=1e-8, rtol=1e-6)
    b = torch.tensor(rng.standard_normal(4), device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(pmcmc_matmul, (A, b), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_matmul_reverse_rule_analytic():
    import torch

    from ParallelMCMC import pmcmc_matmul

    rng = np.random.default_rng(6)
    A = torch.tensor(rng.standard_normal((70, 33)), device="cuda", requires_grad=True)
    B = torch.tensor(rng.standard_normal((33, 90)), device="cuda", requires_grad=True)
    Yb = rng.standard_normal((70, 90))
    pmcmc_matmul(A, B).backward(torch.tensor(Yb, device="cuda"))
    An, Bn = A.detach().cpu().numpy(), B.detach().cpu().numpy()
    assert np.allclose(A.grad.cpu().numpy(), Yb @ Bn.T, rtol=1e-12, atol=1e-12)
    assert np.allclose(B.grad.cpu().numpy(), An.T @ Yb, rtol=1e-12, atol=1e-12)


def test_matmul_forward_rule():
    """Forward JVP dY = dA B + A dB (the reference's forward-mode test)."""
    import torch
    import torch.autograd.forward_ad as fwAD

    from ParallelMCMC import pmcmc_matmul

    rng = np.random.default_rng(1)
    M, K, N = 5, 4, 3
    A, B, dA, dB = (rng.standard_normal(s) for s in [(M, K), (K, N), (M, K), (K, N)])
    with fwAD.dual_level():
        Y = pmcmc_matmul(fwAD.make_dual(_t(A), _t(dA)), fwAD.make_dual(_t(B), _t(dB)))
        dY = fwAD.unpack_dual(Y).tangent.cpu().numpy()
    assert np.allclose(dY, dA @ B + A @ dB, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", [1, 3, 1000, 300001])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_dot_values(n, dt):
    from ParallelMCMC import pmcmc_dot

    rng = np.random.default_rng(n)
    a = rng.standard_normal(n).astype(dt)
    b = rng.standard_normal(n).astype(dt)
    got = float(pmcmc_dot(_t(a), _t(b)).cpu())
    ref = float(a.astype(np.float64) @ b.astype(np.float64))
    # FP64 accumulation: FP64 operands to ~n eps64, FP32 ones to one FP32 rounding of the result
    tol = 1e-12 * np.sum(np.abs(a * b)) if dt == np.float64 else 1e-6 * (abs(ref) + 1e-3 * np.sqrt(n))
    assert abs(got - ref) <= tol


def test_dotsum_values_and_rules():
    import torch
    import torch.autograd.forward_ad as fwAD

    from ParallelMCMC import DimensionMismatch, pmcmc_dotsum

    rng = np.random.default_rng(9)
    A = rng.standard_normal((37, 11))
    B = rng.standard_normal((37, 11))
    assert abs(float(pmcmc_dotsum(_t(A), _t(B)).cpu()) - np.sum(A * B)) < 1e-12
    tA = torch.tensor(A[:6, :4].copy(), device="cuda", requires_grad=True)
    tB = torch.tensor(B[:6, :4].copy(), device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(pmcmc_dotsum, (tA, tB), eps=1e-6, atol=1e-8, rtol=1e-6)
    dA = rng.standard_normal((37, 11))
    dB = rng.standard_normal((37, 11))
    with fwAD.dual_level():
        y = pmcmc_dotsum(fwAD.make_dual(_t(A), _t(dA)), fwAD.make_dual(_t(B), _t(dB)))
        dy = float(fwAD.unpack_dual(y).tangent.cpu())
    assert abs(dy - (np.sum(dA * B) + np.sum(A * dB))) < 1e-11
    with pytest.raises(DimensionMismatch):
        pmcmc_dotsum(_t(A), _t(B[:, :3].copy()))
