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
    """Bayesian logistic regression, prior N(0, prior_scale^2 I), or
    N(0, P^-1) with a dense ``prior_precision`` P (the reparametrized target
    of a mass-matrix preconditioner).
    benchmarks/ParallelMCMCBenchmarks/src/models/bayes_logreg.jl:91-206."""

    def __init__(self, X, y, prior_scale=1.0, prior_precision=None):
        super().__init__()
        self.X = np.asarray(X, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        if self.X.ndim != 2 or self.X.shape[0] != len(self.y):
            raise _lib.DimensionMismatch("X must be N x D and y must have N entries")
        self.N, self.dim = self.X.shape
        self.prior_scale = float(prior_scale)
        self.prior_precision = None
        if prior_precision is not None:
            P = np.asarray(prior_precision, dtype=np.float64)
            if P.shape != (self.dim, self.dim):
                raise _lib.DimensionMismatch("prior_precision must be D x D")
            self.prior_precision = 0.5 * (P + P.T)

    def _create(self, ctx):
        dev = torch.device("cuda", ctx.device)
        Xd = _device.to_colmajor(self.X, torch.float64, dev)
        y = torch.from_numpy(self.y).to(dev)
        h = C.c_void_p()
        ctx.call("pmcmc_target_logistic", self.N, self.dim, _lib.PREC_F64, _device.ptr(Xd), _device.ptr(y),
                 self.prior_scale, C.byref(h))
        if self.prior_precision is not None:
            P = torch.from_numpy(np.ascontiguousarray(self.prior_precision)).to(dev)
            torch.cuda.current_stream(dev).synchronize()
            rc = _lib.load().pmcmc_target_logistic_set_prior(h, _device.ptr(P), _lib.PREC_F64)
            _lib.check(rc, ctx.h, "pmcmc_target_logistic_set_prior")
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


def reparametrized(target, L):
    """The target of theta = L^-1 x for a lower-triangular L with L L' = M:
    p_theta(theta) = p(L theta) (the same log-density values).  Standard MALA
    on theta with the tape (xi, u) maps through x = L theta onto MALA with the
    mass matrix M on x with the same tape (src/MALA/MALA.jl:29-47, 430-438):
    the proposal x + eps M g + sqrt(2 eps) L xi and the acceptance ratio are
    identical.  Gaussian targets stay dense Gaussians (precision L' Lam L,
    mean L^-1 mu); the logistic target gets the design X L and the prior
    precision L' P L."""
    L = np.asarray(L, dtype=np.float64)
    D = target.dim
    if L.shape != (D, D):
        raise _lib.DimensionMismatch("cholM must be D x D")

    def solve_l(v):
        return np.linalg.solve(L, np.asarray(v, dtype=np.float64))

    if isinstance(target, LogisticTarget):
        P = target.prior_precision if target.prior_precision is not None else np.eye(D) / target.prior_scale**2
        return LogisticTarget(target.X @ L, target.y, target.prior_scale, prior_precision=L.T @ P @ L)
    if isinstance(target, GaussianTarget):
        mu = None if target.mean is None else solve_l(target.mean)
        return GaussianTarget(L.T @ target.Lam @ L, mu)
    if isinstance(target, DiagGaussianTarget):
        prec = np.asarray(target.prec, dtype=np.float64)
        mu = None if target.mean is None else solve_l(target.mean)
        return GaussianTarget(L.T @ (prec[:, None] * L), mu)
    raise _lib.ArgumentError(f"no mass-matrix form for {type(target).__name__}")
