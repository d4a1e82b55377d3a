"""Device plumbing: one library context per HIP device, and column-major
device matrices held in torch tensors.

A ``D x T`` trajectory is a torch tensor of shape ``(D, T)`` with strides
``(1, D)`` (Julia's column-major layout), so ``S[:, t]`` is column ``t`` and
``S.data_ptr()`` is what the C ABI receives.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np
import torch

from . import _lib

_ctx = {}
_ctx_lock = threading.Lock()


class Context:
    def __init__(self, device: int):
        lib = _lib.load()
        if not torch.cuda.is_available():
            raise _lib.PmcmcError("no HIP device visible: the DEER kernels need an MI355X (gfx950)")
        self.device = device
        h = C.c_void_p()
        rc = lib.pmcmc_ctx_create(device, None, C.byref(h))
        _lib.check(rc, None, "pmcmc_ctx_create")
        self.h = h

    def call(self, name, *args):
        # inputs may have been produced on torch's stream: make them visible
        torch.cuda.current_stream(self.device).synchronize()
        rc = getattr(_lib.load(), name)(self.h, *args)
        _lib.check(rc, self.h, name)

    def set_profiling(self, on=True, kinds=None):
        """Per-kernel event timing of DEER solves: every kind, or only the names
        in `kinds` (entries of ``_lib.KT_NAMES``)."""
        if kinds is None or not on:
            _lib.check(_lib.load().pmcmc_ctx_set_profiling(self.h, 1 if on else 0), self.h)
            return
        mask = 0
        for name in kinds:
            mask |= 1 << _lib.KT_NAMES.index(name)
        _lib.check(_lib.load().pmcmc_ctx_set_profiling_kinds(self.h, mask), self.h)

    def reset_kernel_times(self):
        _lib.check(_lib.load().pmcmc_ctx_reset_kernel_times(self.h), self.h)

    def kernel_times(self):
        """{name: (launches, total_ms)} of the profiled kernels."""
        out = {}
        for k, name in enumerate(_lib.KT_NAMES):
            n = C.c_int64()
            ms = C.c_double()
            _lib.check(_lib.load().pmcmc_ctx_kernel_time(self.h, k, C.byref(n), C.byref(ms)), self.h)
            out[name] = (int(n.value), float(ms.value))
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().pmcmc_ctx_destroy(self.h)
        except Exception:
            pass


def context(device=None) -> Context:
    if device is None:
        device = torch.cuda.current_device() if torch.cuda.is_available() else 0
    with _ctx_lock:
        if device not in _ctx:
            _ctx[device] = Context(device)
        return _ctx[device]


def torch_dtype(prec):
    return torch.float64 if prec == _lib.PREC_F64 else torch.float32


def np_dtype(prec):
    return np.float64 if prec == _lib.PREC_F64 else np.float32


def colmajor_empty(D, T, dtype, device):
    return torch.empty((T, D), dtype=dtype, device=device).t()


def to_colmajor(M, dtype, device):
    """Device copy of a (D, T) array (numpy or torch) in column-major layout."""
    if isinstance(M, torch.Tensor):
        M = M.detach()
        out = colmajor_empty(M.shape[0], M.shape[1], dtype, device)
        out.copy_(M.to(device=device, dtype=dtype))
        return out
    M = np.asarray(M)
    return torch.from_numpy(np.ascontiguousarray(M.T.astype(np_dtype_of(dtype)))).to(device).t()


def np_dtype_of(tdtype):
    return np.float64 if tdtype == torch.float64 else np.float32


def to_vec(v, dtype, device):
    if isinstance(v, torch.Tensor):
        return v.detach().to(device=device, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(v, dtype=np_dtype_of(dtype)))).to(device)


def is_colmajor(M):
    return M.dim() == 2 and (M.shape[1] <= 1 or M.stride(1) == M.shape[0]) and (M.shape[0] <= 1 or M.stride(0) == 1)


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None
