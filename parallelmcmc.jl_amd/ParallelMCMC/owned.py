"""The owned matmul and reductions (src/ParallelMCMC.jl:23-25) on the device:
``pmcmc_matmul(A, B) = A * B``, ``pmcmc_dot(a, b) = dot(a, b)``,
``pmcmc_dotsum(A, B) = sum(A .* B)``, computed by the library's MFMA GEMM and
fixed-order reductions (``pmcmc_matmul`` / ``pmcmc_dot`` / ``pmcmc_dotsum``
of the C ABI).

They carry the derivative rules of the reference's AD extension
(ext/EnzymeExt.jl:78-284) as torch autograd rules:
* reverse: matmul dA = dY B^T, dB = A^T dY (transposed products through the
  same GEMM); dot / dotsum dA = dy B, dB = dy A;
* forward: matmul dA B + A dB; dot / dotsum <dA, B> + <A, dB>.

Operands are CUDA (HIP) torch tensors of float32 or float64, in torch's
row-major layout; the column-major ABI sees a row-major r x c buffer as the
c x r matrix, so op(X) op(Y) is computed as (op(Y)^T op(X)^T)^T with the
transpose flags swapped.  Vectors follow Julia: ``A * b`` with a vector b is a
vector, ``a * B`` with a vector a needs B to have one row.
"""
from __future__ import annotations

import torch

from . import _device, _lib
from ._lib import ArgumentError, DimensionMismatch


def _prec(*ts):
    dt = ts[0].dtype
    if any(t.dtype != dt for t in ts):
        raise ArgumentError("operands must have the same element type")
    if dt == torch.float64:
        return _lib.PREC_F64
    if dt == torch.float32:
        return _lib.PREC_F32
    raise ArgumentError(f"element type {dt} is not supported (float32 or float64)")


def _device_of(*ts):
    dev = ts[0].device
    if dev.type != "cuda" or any(t.device != dev for t in ts):
        raise ArgumentError("operands must be device tensors on the same HIP device")
    return dev


def _gemm(X, Y, tx=False, ty=False):
    """op(X) @ op(Y) for 2-D row-major tensors (op = transpose when flagged)."""
    rows = X.shape[1] if tx else X.shape[0]
    k = X.shape[0] if tx else X.shape[1]
    k2 = Y.shape[1] if ty else Y.shape[0]
    cols = Y.shape[0] if ty else Y.shape[1]
    if k != k2:
        raise DimensionMismatch(f"matrix A has dimensions {tuple(X.shape)}, matrix B has dimensions {tuple(Y.shape)}")
    prec = _prec(X, Y)
    dev = _device_of(X, Y)
    X = X.contiguous()
    Y = Y.contiguous()
    Z = torch.empty(rows, cols, dtype=X.dtype, device=dev)
    if rows == 0 or cols == 0:
        return Z
    ctx = _device.context(dev.index)
    ctx.call("pmcmc_matmul", prec, 1 if ty else 0, 1 if tx else 0, cols, rows, k, _device.ptr(Y),
             max(1, Y.shape[1]), _device.ptr(X), max(1, X.shape[1]), _device.ptr(Z), max(1, cols))
    return Z


def _dotsum_raw(A, B):
    if A.shape != B.shape:
        raise DimensionMismatch(f"arrays could not be broadcast to a common size; got {tuple(A.shape)} and "
                                f"{tuple(B.shape)}")
    prec = _prec(A, B)
    dev = _device_of(A, B)
    A = A.contiguous()
    B = B.contiguous()
    out = torch.empty((), dtype=A.dtype, device=dev)
    n = A.numel()
    ctx = _device.context(dev.index)
    ctx.call("pmcmc_dotsum", prec, n, 1, _device.ptr(A), max(1, n), _device.ptr(B), max(1, n), _device.ptr(out))
    return out


class _MatMul(torch.autograd.Function):
    @staticmethod
    def forward(A, B):
        return _gemm(A, B)

    @staticmethod
    def setup_context(ctx, inputs, output):
        A, B = inputs
        ctx.save_for_backward(A, B)
        ctx.save_for_forward(A, B)

    @staticmethod
    def backward(ctx, gY):
        A, B = ctx.saved_tensors
        gA = _gemm(gY, B, ty=True) if ctx.needs_input_grad[0] else None
        gB = _gemm(A, gY, tx=True) if ctx.needs_input_grad[1] else None
        return gA, gB

    @staticmethod
    def jvp(ctx, dA, dB):
        A, B = ctx.saved_tensors
        out = _gemm(dA, B) if dA is not None else None
        if dB is not None:
            t = _gemm(A, dB)
            out = t if out is None else out + t
        return out


class _DotSum(torch.autograd.Function):
    @staticmethod
    def forward(A, B):
        return _dotsum_raw(A, B)

    @staticmethod
    def setup_context(ctx, inputs, output):
        A, B = inputs
        ctx.save_for_backward(A, B)
        ctx.save_for_forward(A, B)

    @staticmethod
    def backward(ctx, gy):
        A, B = ctx.saved_tensors
        return (gy * B if ctx.needs_input_grad[0] else None, gy * A if ctx.needs_input_grad[1] else None)

    @staticmethod
    def jvp(ctx, dA, dB):
        A, B = ctx.saved_tensors
        out = _dotsum_raw(dA, B) if dA is not None else None
        if dB is not None:
            t = _dotsum_raw(A, dB)
            out = t if out is None else out + t
        return out


def pmcmc_matmul(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """``A * B`` (src/ParallelMCMC.jl:23) for 2-D / 1-D device tensors."""
    if A.dim() == 1 and B.dim() == 1:
        raise ArgumentError("vector * vector is undefined; use pmcmc_dot")
    if A.dim() not in (1, 2) or B.dim() not in (1, 2):
        raise ArgumentError("operands must be vectors or matrices")
    vec_out = B.dim() == 1
    A2 = A.unsqueeze(1) if A.dim() == 1 else A
    B2 = B.unsqueeze(1) if B.dim() == 1 else B
    Y = _MatMul.apply(A2, B2)
    return Y.squeeze(1) if vec_out else Y


def pmcmc_dot(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """``dot(a, b)`` (src/ParallelMCMC.jl:24) for device vectors; a 0-d tensor."""
    if a.dim() != 1 or b.dim() != 1:
        raise ArgumentError("pmcmc_dot takes two vectors")
    if a.shape[0] != b.shape[0]:
        raise DimensionMismatch(f"first array has length {a.shape[0]} which does not match the length of the "
                                f"second, {b.shape[0]}")
    return _DotSum.apply(a, b)


def pmcmc_dotsum(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """``sum(A .* B)`` (src/ParallelMCMC.jl:25) for device arrays of one shape; a 0-d tensor."""
    return _DotSum.apply(A, B)
