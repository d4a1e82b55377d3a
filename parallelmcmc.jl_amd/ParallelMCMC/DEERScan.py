"""Diagonal affine scan on the GPU.  Mirrors src/DEER/DEERScan.jl.

    s_t = A[:, t] .* s_{t-1} + B[:, t],   t = 1..T,  s_0 = s0

computed by one single-pass HIP kernel (LDS-staged tiles + decoupled
look-back) instead of the reference's log2(T) whole-array doubling passes.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _device, _lib
from ._lib import ArgumentError, DimensionMismatch


class AffineScanWorkspace:
    """Shape/dtype record; the device scratch lives in the library context
    (src/DEER/DEERScan.jl:18-31)."""

    def __init__(self, A_template):
        self.shape = tuple(A_template.shape)
        self.dtype = A_template.dtype


def _as_device(M, dtype, device, vec=False):
    if vec:
        return _device.to_vec(M, dtype, device)
    if isinstance(M, torch.Tensor) and M.device == device and M.dtype == dtype and _device.is_colmajor(M):
        return M
    return _device.to_colmajor(M, dtype, device)


def solve_affine_scan_diag_(S, A, B, s0, ws=None, check_shapes=True):
    """In-place form (``solve_affine_scan_diag!``, src/DEER/DEERScan.jl:102-173).

    ``S``, ``A``, ``B`` are column-major device tensors of shape (D, T)
    (float32 or float64), ``s0`` a length-D vector.  Returns ``S``."""
    D, T = A.shape
    if check_shapes:
        if tuple(B.shape) != (D, T):
            raise DimensionMismatch("B must have the same size as A")
        if tuple(S.shape) != (D, T):
            raise DimensionMismatch("S must have size (D, T)")
        if len(s0) != D:
            raise DimensionMismatch("length(s0) must equal size(A,1)")
        if ws is not None and tuple(ws.shape) != (D, T):
            raise DimensionMismatch("workspace has wrong matrix size")
    if not isinstance(S, torch.Tensor) or S.device.type != "cuda" or not _device.is_colmajor(S):
        raise ArgumentError("S must be a column-major device tensor (shape (D,T), strides (1,D))")
    dtype = S.dtype
    if dtype not in (torch.float32, torch.float64):
        raise ArgumentError("S must be float32 or float64")
    dev = S.device
    A_ = _as_device(A, dtype, dev)
    B_ = _as_device(B, dtype, dev)
    s0_ = _as_device(s0, dtype, dev, vec=True)
    ctx = _device.context(dev.index)
    name = "pmcmc_scan_diag_f64" if dtype == torch.float64 else "pmcmc_scan_diag_f32"
    ctx.call(name, _device.ptr(A_), _device.ptr(B_), _device.ptr(s0_), _device.ptr(S), D, T)
    return S


def solve_affine_scan_diag(A, B, s0, device=None):
    """Allocating form (src/DEER/DEERScan.jl:186-191)."""
    if isinstance(A, torch.Tensor):
        dtype, dev = A.dtype, A.device
    else:
        dtype = torch.float64 if np.asarray(A).dtype == np.float64 else torch.float32
        dev = torch.device("cuda", device if device is not None else torch.cuda.current_device())
    D, T = A.shape
    S = _device.colmajor_empty(D, T, dtype, dev)
    return solve_affine_scan_diag_(S, A, B, s0)
