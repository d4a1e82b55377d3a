"""Affine scans on the GPU.  Mirrors src/DEER/DEERScan.jl.

    s_t = A[:, t] .* s_{t-1} + B[:, t],   t = 1..T,  s_0 = s0

computed by LDS-tiled reduce -> spine -> apply kernels (a single pass with
decoupled look-back is available through PMCMC_SCAN_CHAINED=1) instead of the
reference's log2(T) whole-array doubling passes; and the dense form

    s_t = J_t s_{t-1} + B[:, t]

with one D x D Jacobian per column (D <= 64), which the reference does not
have (it rejects ``:full``).
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


def solve_affine_scan_full(J, B, s0, device=None):
    """Dense-Jacobian scan ``s_t = J[:, :, t] s_{t-1} + B[:, t]`` (D <= 64).

    ``J`` is (D, D, T) and ``B`` (D, T); numpy arrays or device tensors
    (copied to column-major device memory when not already so).  Returns the
    (D, T) column-major device tensor S."""
    if isinstance(J, torch.Tensor):
        dtype, dev = J.dtype, J.device
        if dev.type != "cuda":
            dev = torch.device("cuda", device if device is not None else torch.cuda.current_device())
    else:
        dtype = torch.float64 if np.asarray(J).dtype == np.float64 else torch.float32
        dev = torch.device("cuda", device if device is not None else torch.cuda.current_device())
    if dtype not in (torch.float32, torch.float64):
        raise ArgumentError("J must be float32 or float64")
    if len(J.shape) != 3 or J.shape[0] != J.shape[1]:
        raise DimensionMismatch("J must have size (D, D, T)")
    D, _, T = J.shape
    if tuple(B.shape) != (D, T):
        raise DimensionMismatch("B must have size (D, T)")
    if len(s0) != D:
        raise DimensionMismatch("length(s0) must equal size(J,1)")
    if D > 64:
        raise ArgumentError("the dense scan supports D <= 64")
    if isinstance(J, torch.Tensor):
        Jc = J.to(device=dev, dtype=dtype).permute(2, 1, 0).contiguous()  # (T, k, i): column-major per matrix
    else:
        Jc = torch.tensor(np.ascontiguousarray(np.transpose(np.asarray(J), (2, 1, 0))), dtype=dtype, device=dev)
    B_ = _as_device(B, dtype, dev)
    s0_ = _as_device(s0, dtype, dev, vec=True)
    S = _device.colmajor_empty(D, T, dtype, dev)
    ctx = _device.context(dev.index)
    name = "pmcmc_scan_full_f64" if dtype == torch.float64 else "pmcmc_scan_full_f32"
    ctx.call(name, _device.ptr(Jc), _device.ptr(B_), _device.ptr(s0_), _device.ptr(S), D, T)
    return S
