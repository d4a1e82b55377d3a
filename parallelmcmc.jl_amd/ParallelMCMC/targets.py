"""GPU target descriptors.

The reference's DensityModel carries user closures (logdensity, grad, hvp and
their batched forms, src/interface.jl:36-70).  Closures cannot run inside HIP
kernels, so a DensityModel that should run on the GPU carries one of these
analytic descriptors; the kernels evaluate the same functions in closed form.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _device, _lib


class GPUTarget:
    dim: int

    def __init__(self):
        self._handles = {}

    def handle(self, ctx: _device.Context):
        key = ctx.device
        if key not in self._handles:
            self._handles[key] = (ctx, self._create(ctx))
        return self._handles[key][1]

    def __del__(self):
        for ctx, h in getattr(self, "_handles", {}).values():
            try:
                _lib.load().pmcmc_target_destroy(h)
            except Exception:
                pass


class DiagGaussianTarget(GPUTarget):
    """N(mean, diag(1/prec)); logp = -1/2 sum prec (x - mean)^2."""

    def __init__(self, prec, mean=None):
        super().__init__()
        self.prec = np.asarray(prec, dtype=np.float64)
        self.mean = None if mean is None else np.asarray(mean, dtype=np.float64)
        self.dim = len(self.prec)

    def _create(self, ctx):
        dev = torch.device("cuda", ctx.device)
        p = torch.from_numpy(self.prec).to(dev)
        m = torch.from_numpy(self.mean).to(dev) if self.mean is not None else None
        h = C.c_void_p()
        ctx.call("pmcmc_target_diag_gaussian", self.dim, _lib.PREC_F64, _device.ptr(p), _device.ptr(m), C.byref(h))
        return h


class StdNormalTarget(DiagGaussianTarget):
    """logp = -||x||^2 / 2, grad = -x, Hv = -v (test/test-Deer-vs-MALA.jl:13-14)."""

    def __init__(self, dim):
        GPUTarget.__init__(self)
        self.prec = np.ones(dim)
        self.mean = None
        self.dim = dim

    def _create(self, ctx):
        h = C.c_void_p()
        ctx.call("pmcmc_target_diag_gaussian", self.dim, _lib.PREC_F64, None, None, C.byref(h))
        return h


class LogisticTarget(GPUTarget):
    """Bayesian logistic regression, prior N(0, prior_scale^2 I).
    benchmarks/ParallelMCMCBenchmarks/src/models/bayes_logreg.jl:91-206."""

    def __init__(self, X, y, prior_scale=1.0):
        super().__init__()
        self.X = np.asarray(X, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        if self.X.ndim != 2 or self.X.shape[0] != len(self.y):
            raise _lib.DimensionMismatch("X must be N x D and y must have N entries")
        self.N, self.dim = self.X.shape
        self.prior_scale = float(prior_scale)

    def _create(self, ctx):
        dev = torch.device("cuda", ctx.device)
        Xd = _device.to_colmajor(self.X, torch.float64, dev)
        y = torch.from_numpy(self.y).to(dev)
        h = C.c_void_p()
        ctx.call("pmcmc_target_logistic", self.N, self.dim, _lib.PREC_F64, _device.ptr(Xd), _device.ptr(y),
                 self.prior_scale, C.byref(h))
        return h


class GaussianTarget(GPUTarget):
    """N(mean, Lam^-1) with a dense symmetric precision Lam: grad = -Lam (x - mean),
    Hv = -Lam v; evaluated as MFMA GEMMs over all trajectory columns."""

    def __init__(self, Lam, mean=None):
        super().__init__()
        self.Lam = np.asarray(Lam, dtype=np.float64)
        if self.Lam.ndim != 2 or self.Lam.shape[0] != self.Lam.shape[1]:
            raise _lib.DimensionMismatch("Lam must be square")
        self.mean = None if mean is None else np.asarray(mean, dtype=np.float64)
        self.dim = self.Lam.shape[0]

    def _create(self, ctx):
        dev = torch.device("cuda", ctx.device)
        L = _device.to_colmajor(self.Lam, torch.float64, dev)
        m = torch.from_numpy(self.mean).to(dev) if self.mean is not None else None
        h = C.c_void_p()
        ctx.call("pmcmc_target_gaussian", self.dim, _lib.PREC_F64, _device.ptr(L), _device.ptr(m), C.byref(h))
        return h
