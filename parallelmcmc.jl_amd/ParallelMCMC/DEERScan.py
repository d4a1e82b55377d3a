"""Affine scans on the GPU.  Mirrors src/DEER/DEERScan.jl (exports :3-9):
``solve_affine_scan_diag(!)``, ``solve_affine_seq(!)``,
``affine_scan_residual``, ``check_affine_scan``, ``AffineScanWorkspace``.

    s_t = A[:, t] .* s_{t-1} + B[:, t],   t = 1..T,  s_0 = s0

computed by LDS-tiled reduce -> spine -> apply kernels (a single pass with
decoupled look-back is available through PMCMC_SCAN_CHAINED=1) instead of the
reference's log2(T) whole-array doubling passes; and the dense form

    s_t = J_t s_{t-1} + B[:, t]

with one D x D Jacobian per column (D <= 64), which the reference does not
have (it rejects ``:full``).
"""
from __future__ import annotations

import ctypes as C

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


def _dev_dtype(A, device):
    if isinstance(A, torch.Tensor) and A.device.type == "cuda":
        dtype, dev = A.dtype, A.device
    else:
        a = A.cpu().numpy() if isinstance(A, torch.Tensor) else np.asarray(A)
        dtype = torch.float64 if a.dtype == np.float64 else torch.float32
        dev = torch.device("cuda", device if device is not None else torch.cuda.current_device())
    if dtype not in (torch.float32, torch.float64):
        raise ArgumentError("A must be float32 or float64")
    return dtype, dev


def _check_dims(A, B, s0, S=None):
    D, T = A.shape
    if tuple(B.shape) != (D, T):
        raise DimensionMismatch("B must have the same size as A")
    if S is not None and tuple(S.shape) != (D, T):
        raise DimensionMismatch("S must have size (D, T)")
    if len(s0) != D:
        raise DimensionMismatch("length(s0) must equal size(A,1)")
    return D, T


def solve_affine_seq_(S, A, B, s0):
    """``solve_affine_seq!`` (src/DEER/DEERScan.jl:51-68): the sequential
    recurrence on the device, one thread per row, product and sum rounded
    separately.  ``S`` is a column-major (D, T) device tensor; returns it."""
    D, T = _check_dims(A, B, s0, S)
    if not isinstance(S, torch.Tensor) or S.device.type != "cuda" or not _device.is_colmajor(S):
        raise ArgumentError("S must be a column-major device tensor (shape (D,T), strides (1,D))")
    dtype, dev = S.dtype, S.device
    if dtype not in (torch.float32, torch.float64):
        raise ArgumentError("S must be float32 or float64")
    if T == 0:
        return S
    A_, B_ = _as_device(A, dtype, dev), _as_device(B, dtype, dev)
    s0_ = _as_device(s0, dtype, dev, vec=True)
    name = "pmcmc_affine_seq_f64" if dtype == torch.float64 else "pmcmc_affine_seq_f32"
    _device.context(dev.index).call(name, _device.ptr(A_), _device.ptr(B_), _device.ptr(s0_), _device.ptr(S), D, T)
    return S


def solve_affine_seq(A, B, s0, device=None):
    """Allocating form of ``solve_affine_seq!`` (src/DEER/DEERScan.jl:70-75)."""
    D, T = _check_dims(A, B, s0)
    dtype, dev = _dev_dtype(A, device)
    return solve_affine_seq_(_device.colmajor_empty(D, T, dtype, dev), A, B, s0)


def affine_scan_residual(S, A, B, s0, device=None):
    """Maximum absolute residual of the recurrence (src/DEER/DEERScan.jl:
    203-224): max |S[:, t] - (A[:, t] .* S[:, t-1] + B[:, t])| with S[:, 0]
    from s0; 0 for T == 0, NaN when any entry is NaN.  Computed on the
    device; returns a float."""
    D, T = _check_dims(A, B, s0, S)
    dtype, dev = _dev_dtype(S, device)
    if T == 0:
        return 0.0
    S_, A_, B_ = (_as_device(M, dtype, dev) for M in (S, A, B))
    s0_ = _as_device(s0, dtype, dev, vec=True)
    out = C.c_double()
    name = "pmcmc_affine_residual_f64" if dtype == torch.float64 else "pmcmc_affine_residual_f32"
    _device.context(dev.index).call(name, _device.ptr(S_), _device.ptr(A_), _device.ptr(B_), _device.ptr(s0_), D, T,
                                    C.byref(out))
    return float(out.value)


def check_affine_scan(A, B, s0, atol=1e-6, rtol=1e-6, device=None):
    """Sequential solver against the parallel scan (src/DEER/DEERScan.jl:
    240-258).  Returns a dict with ok, max_abs_err, max_rel_err,
    residual_seq, residual_scan, S_seq, S_scan (device tensors)."""
    D, T = _check_dims(A, B, s0)
    S_seq = solve_affine_seq(A, B, s0, device)
    S_scan = solve_affine_scan_diag(A, B, s0, device)
    dtype, dev = S_seq.dtype, S_seq.device
    out = (C.c_double * 3)()
    name = "pmcmc_affine_compare_f64" if dtype == torch.float64 else "pmcmc_affine_compare_f32"
    _device.context(dev.index).call(name, _device.ptr(S_seq), _device.ptr(S_scan), D * T, float(atol), float(rtol),
                                    out)
    max_abs, max_rel = float(out[0]), float(out[1])
    return {
        "ok": bool(out[2] == 1.0),
        "max_abs_err": max_abs,
        "max_rel_err": max_rel,
        "residual_seq": affine_scan_residual(S_seq, A, B, s0),
        "residual_scan": affine_scan_residual(S_scan, A, B, s0),
        "S_seq": S_seq,
        "S_scan": S_scan,
    }


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
