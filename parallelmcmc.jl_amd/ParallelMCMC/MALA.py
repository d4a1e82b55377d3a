"""Batched taped MALA step + surrogate JVP on the GPU.

Mirrors MALA.mala_step_batched_fwd_and_jvp! (src/MALA/MALA.jl:556-628): for
every column t, the exact taped MALA step from X[:,t] (noise Xi[:,t], uniform
U[t]) and the JVP of the sigmoid surrogate along Z[:,t].  The target's
gradient / HVP / log-density are the GPU target descriptor's closed forms
instead of user closures.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _device, _lib
from ._lib import ArgumentError, DimensionMismatch
from .interface import _precision_of, _target_of


def mala_step_batched_fwd_and_jvp(model, X, eps, Xi, U, Z, precision=None, device=None, return_logalpha=False):
    """Returns (FT, Jt, accepted) as device tensors (accepted is uint8)."""
    tgt = _target_of(model)
    prec = _precision_of(eps, precision)
    D, T = X.shape
    if tuple(Xi.shape) != (D, T):
        raise DimensionMismatch("X and Xi must have the same size")
    if tuple(Z.shape) != (D, T):
        raise DimensionMismatch("X and Z must have the same size")
    if len(U) != T:
        raise DimensionMismatch("u must have length size(X,2)")
    if D != tgt.dim:
        raise DimensionMismatch("X rows must equal the model dimension")
    dev = torch.device("cuda", device if device is not None else torch.cuda.current_device())
    dt = _device.torch_dtype(prec)
    X_, Xi_, Z_ = (_device.to_colmajor(M, dt, dev) for M in (X, Xi, Z))
    U_ = _device.to_vec(U, dt, dev)
    FT = _device.colmajor_empty(D, T, dt, dev)
    Jt = _device.colmajor_empty(D, T, dt, dev)
    acc = torch.empty(T, dtype=torch.uint8, device=dev)
    la = torch.empty(T, dtype=torch.float64, device=dev)
    ctx = _device.context(dev.index)
    ctx.call("pmcmc_mala_fwd_and_jvp", tgt.handle(ctx), prec, float(eps), _device.ptr(X_), _device.ptr(Xi_),
             _device.ptr(U_), _device.ptr(Z_), _device.ptr(FT), _device.ptr(Jt), _device.ptr(acc),
             _device.ptr(la), D, T)
    if return_logalpha:
        return FT, Jt, acc, la
    return FT, Jt, acc


def logdensity_batch(model, S, precision=None, device=None):
    """logp of every column of S as float64 (src/interface.jl:517-524)."""
    tgt = _target_of(model)
    if precision is None:
        is64 = (S.dtype == torch.float64) if isinstance(S, torch.Tensor) else (np.asarray(S).dtype == np.float64)
        prec = _lib.PREC_F64 if is64 else _lib.PREC_F32
    else:
        prec = _precision_of(1.0, precision)
    D, T = S.shape
    if D != tgt.dim:
        raise DimensionMismatch("S rows must equal the model dimension")
    if isinstance(S, torch.Tensor) and S.device.type == "cuda":
        dev = S.device
    else:
        dev = torch.device("cuda", device if device is not None else torch.cuda.current_device())
    dt = _device.torch_dtype(prec)
    if isinstance(S, torch.Tensor) and S.device == dev and S.dtype == dt and _device.is_colmajor(S):
        S_ = S
    else:
        S_ = _device.to_colmajor(S, dt, dev)
    out = torch.empty(T, dtype=torch.float64, device=dev)
    ctx = _device.context(dev.index)
    ctx.call("pmcmc_target_logdensity_batch", tgt.handle(ctx), prec, _device.ptr(S_), _device.ptr(out), T)
    return out


def mala_run_batched(model, X0, eps, Xi, U, precision=None, device=None):
    """Sequential MALA on the device for N chains of one model, K steps
    (each step is mala_step_batched!, src/MALA/MALA.jl:487-528, on the N
    columns).  X0: (D, N); Xi: (D, N, K); U: (N, K).  Returns (S, accepted):
    S (D, N, K) device tensor, S[:, j, k] the state of chain j after step k;
    accepted (N, K) uint8."""
    tgt = _target_of(model)
    prec = _precision_of(eps, precision)
    X0 = np.asarray(X0) if not isinstance(X0, torch.Tensor) else X0
    if len(X0.shape) != 2:
        raise DimensionMismatch("X0 must be D x N")
    D, N = X0.shape
    if D != tgt.dim:
        raise DimensionMismatch("X0 rows must equal the model dimension")
    if len(Xi.shape) != 3 or tuple(Xi.shape[:2]) != (D, N):
        raise DimensionMismatch("Xi must be D x N x K")
    K = Xi.shape[2]
    if tuple(U.shape) != (N, K):
        raise DimensionMismatch("U must be N x K")
    dev = torch.device("cuda", device if device is not None else torch.cuda.current_device())
    dt = _device.torch_dtype(prec)
    X0_ = _device.to_colmajor(X0, dt, dev)
    # step k's N columns contiguous: (D, N, K) column-major is exactly that
    Xi_ = torch.as_tensor(np.asarray(Xi) if not isinstance(Xi, torch.Tensor) else Xi).to(dev, dt)
    Xi_ = Xi_.permute(2, 1, 0).contiguous()  # (K, N, D): element (d, j, k) at (k N + j) D + d
    U_ = torch.as_tensor(np.asarray(U) if not isinstance(U, torch.Tensor) else U).to(dev, dt)
    U_ = U_.t().contiguous()  # (K, N)
    S = torch.empty((K, N, D), dtype=dt, device=dev)
    acc = torch.empty((K, N), dtype=torch.uint8, device=dev)
    ctx = _device.context(dev.index)
    ctx.call("pmcmc_mala_run_batched", tgt.handle(ctx), prec, float(eps), _device.ptr(X0_), _device.ptr(Xi_),
             _device.ptr(U_), _device.ptr(S), _device.ptr(acc), D, N, K)
    return S.permute(2, 1, 0), acc.t()


def mala_step_batched(model, X, eps, Xi, U, precision=None, device=None):
    """One step of mala_step_batched! (src/MALA/MALA.jl:487-528) for the N
    columns of X: returns (X_next, accepted)."""
    D, N = X.shape
    S, acc = mala_run_batched(model, X, eps, np.asarray(Xi).reshape(D, N, 1) if not isinstance(Xi, torch.Tensor)
                              else Xi.reshape(D, N, 1), np.asarray(U).reshape(N, 1) if not isinstance(U, torch.Tensor)
                              else U.reshape(N, 1), precision, device)
    return S[:, :, 0], acc[:, 0]
