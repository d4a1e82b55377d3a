"""DEER Newton solve over a taped MALA recursion, on the GPU.

Mirrors src/DEER/DEER.jl: ``TapedRecursion`` (:26-38), ``DEERWorkspace``
(:47-90), ``deer_update`` (:375-492) and ``solve`` (:506-595).  The whole
Newton loop runs on the device (fused MALA step + JVP + Hutchinson diagonal
+ single-pass scan + damping + convergence maxima per iteration); the host
reads one convergence word per chunk of iterations.

Probes are drawn on the device from ``seed`` (the reference draws them from
the host ``rng``, :298-325); they change only the iteration count.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _device, _lib
from ._lib import ArgumentError, DimensionMismatch
from .interface import _precision_of, _target_of


@dataclass
class TapedRecursion:
    """The taped MALA recursion s_t = f_t(s_{t-1}) of ``model`` with step
    ``eps`` driven by the tape (Xi, U) (built by _build_mala_deer_rec,
    src/interface.jl:360-473).  Xi is D x T, U has T entries; both may be
    numpy arrays or device tensors."""

    model: object
    eps: float
    Xi: object
    U: object
    precision: int = None

    def __post_init__(self):
        self.precision = _precision_of(self.eps, self.precision)
        self.target = _target_of(self.model)
        D, T = self.Xi.shape
        if len(self.U) != T:
            raise DimensionMismatch("tape_uniforms must have length T")
        if D != self.target.dim:
            raise DimensionMismatch("tape_noise must have size (D, T)")
        self.D, self.T = D, T

    def device_tape(self, dev):
        dt = _device.torch_dtype(self.precision)
        return _device.to_colmajor(self.Xi, dt, dev), _device.to_vec(self.U, dt, dev)


class DEERWorkspace:
    """Device buffers for solves of shape (D, T) at one precision
    (src/DEER/DEER.jl:47-90); reused across solves."""

    def __init__(self, D, T, precision=_lib.PREC_F64, device=None):
        self.D, self.T, self.precision = int(D), int(T), int(precision)
        self.device = torch.device("cuda", device if device is not None else torch.cuda.current_device())
        self.ctx = _device.context(self.device.index)
        h = C.c_void_p()
        self.ctx.call("pmcmc_deer_ws_create", self.D, self.T, self.precision, C.byref(h))
        self.h = h
        dt = _device.torch_dtype(self.precision)
        self.S_tmp = _device.colmajor_empty(self.D, self.T, dt, self.device)
        self.accepted = torch.empty(self.T, dtype=torch.uint8, device=self.device)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().pmcmc_deer_ws_destroy(self.h)
        except Exception:
            pass


@dataclass
class Info:
    converged: bool
    iters: int
    metric: float
    jacobian: str
    damping: float
    probes: int
    accepted: torch.Tensor = None


def _jac_code(jacobian):
    j = str(jacobian).lstrip(":")
    if j == "stoch_diag":
        return _lib.JAC_STOCH_DIAG, j
    if j == "diag":
        return _lib.JAC_DIAG, j
    if j == "full":
        return _lib.JAC_FULL, j
    if j == "full_sigmoid":
        return _lib.JAC_FULL_SIGMOID, j
    raise ArgumentError("jacobian must be :diag, :stoch_diag, :full or :full_sigmoid")


def solve(rec: TapedRecursion, s0, init=None, tol_abs=1e-4, tol_rel=1e-3, maxiter=200,
          jacobian="stoch_diag", damping=1.0, probes=1, seed=0, return_info=False, workspace=None,
          copy_result=None, device=None, fixed_probes=False):
    """Newton iterations to the fixed point of the taped recursion
    (src/DEER/DEER.jl:506-595).  Returns S (D x T device tensor) and, with
    ``return_info``, an Info record whose ``accepted`` is the final
    iteration's accept mask.  ``fixed_probes`` keeps one Rademacher probe set
    for the whole solve instead of drawing new ones every iteration (the
    reference's choice, src/DEER/DEER.jl:407): the Newton iteration is then a
    stationary linear iteration, which on ill-conditioned targets does not
    show the transient growth that can overflow FP32 with redrawn probes.

    ``jacobian``: ``:stoch_diag`` (Hutchinson, ``probes`` Rademacher
    probes), ``:diag`` (the exact diagonal: closed form for diagonal-Hessian
    targets; for the logistic and dense Gaussian targets, D <= 256, two
    Hessian-vector GEMM passes and diag(H) per iteration instead of the
    reference's D basis JVPs, src/DEER/DEER.jl:327-342), ``:full`` /
    ``:full_sigmoid`` (dense Gaussian, D <= 64)."""
    if maxiter < 1:
        raise ArgumentError("maxiter must be >= 1")
    if tol_abs < 0:
        raise ArgumentError("tol_abs must be >= 0")
    if tol_rel < 0:
        raise ArgumentError("tol_rel must be >= 0")
    if not (0 < damping <= 1):
        raise ArgumentError("damping must be in (0,1]")
    if probes < 1:
        raise ArgumentError("probes must be >= 1")
    jcode, jname = _jac_code(jacobian)
    D, T = rec.D, rec.T
    if jname in ("full", "full_sigmoid"):
        from .targets import GaussianTarget

        if not isinstance(rec.target, GaussianTarget):
            raise ArgumentError("jacobian=:full needs a GaussianTarget (dense precision matrix)")
        if D > 64:
            raise ArgumentError("jacobian=:full supports D <= 64")
    if len(s0) != D:
        raise ArgumentError("S must be D x T with D = length(s0)")
    if copy_result is None:
        copy_result = workspace is not None
    ws = workspace if workspace is not None else DEERWorkspace(D, T, rec.precision, device)
    if (ws.D, ws.T) != (D, T):
        raise ArgumentError("workspace state buffer has wrong size")
    if ws.precision != rec.precision:
        raise ArgumentError("workspace precision does not match the recursion")
    dev = ws.device
    dt = _device.torch_dtype(rec.precision)
    Xi, U = rec.device_tape(dev)
    x0 = _device.to_vec(s0, dt, dev)
    init_ = None
    if init is not None:
        if tuple(init.shape) != (D, T):
            raise ArgumentError("init must be size (D,T)")
        init_ = _device.to_colmajor(init, dt, dev)
    params = _lib.DeerParams(float(rec.eps), float(tol_abs), float(tol_rel), float(damping), int(maxiter),
                             jcode, int(probes), rec.precision, int(seed) & 0xFFFFFFFFFFFFFFFF,
                             1 if fixed_probes else 0)
    info = _lib.DeerInfo()
    ws.ctx.call("pmcmc_deer_solve", ws.h, rec.target.handle(ws.ctx), C.byref(params), _device.ptr(x0),
                _device.ptr(Xi), _device.ptr(U), _device.ptr(init_), _device.ptr(ws.S_tmp),
                _device.ptr(ws.accepted), C.byref(info))
    S = ws.S_tmp.t().clone().t() if copy_result else ws.S_tmp
    if not return_info:
        return S
    acc = ws.accepted.clone() if copy_result else ws.accepted
    return S, Info(bool(info.converged), int(info.iters), float(info.metric), jname, float(damping), int(probes), acc)


def last_trajectory(workspace):
    """The trajectory and its log-densities from the final Newton iteration
    of the workspace's last solve on a logistic or dense Gaussian target,
    assembled on the device: S = FT(S_k), the taped step applied to that
    iteration's iterate (within the solve's tolerance of the returned
    iterate), and logp(S[:, t]) as that iteration's GEMM passes evaluated it
    -- no separate log-density pass (src/interface.jl:517-524).  Returns
    (S, logps): a new D x T device tensor and T float64 device tensor."""
    ws = workspace
    dt = _device.torch_dtype(ws.precision)
    S = _device.colmajor_empty(ws.D, ws.T, dt, ws.device)
    lp = torch.empty(ws.T, dtype=torch.float64, device=ws.device)
    rc = _lib.load().pmcmc_deer_last_trajectory(ws.h, _device.ptr(S), _device.ptr(lp))
    _lib.check(rc, ws.ctx.h, "pmcmc_deer_last_trajectory")
    return S, lp


def adjoint(rec: TapedRecursion, s0, S, accepted, G, tol_abs=0.0, tol_rel=None, maxiter=500, seed=0,
            workspace=None, return_lambda=False):
    """Gradient of a statistic L(S) of the converged trajectory S of ``rec``
    with the accept decisions ``accepted`` held fixed, given G = dL/dS.

    Returns (dL/deps, dL/dhyper, dL/ds0, info[, lambda]); dL/dhyper is the
    derivative with respect to the logistic target's prior scale (0 for the
    Gaussian targets).  No reference counterpart (the reference differentiates
    sequential MALA with Enzyme); the reverse recurrence is solved by
    reverse-time DEER sweeps to tol_abs / tol_rel (default 1e-6 in FP32,
    1e-12 in FP64)."""
    D, T = rec.D, rec.T
    if tol_rel is None:
        tol_rel = 1e-12 if rec.precision == _lib.PREC_F64 else 1e-6
    if maxiter < 1:
        raise ArgumentError("maxiter must be >= 1")
    if tol_abs < 0 or tol_rel < 0:
        raise ArgumentError("tolerances must be >= 0")
    ws = workspace if workspace is not None else DEERWorkspace(D, T, rec.precision)
    if (ws.D, ws.T) != (D, T) or ws.precision != rec.precision:
        raise ArgumentError("workspace does not match the recursion")
    dev = ws.device
    dt = _device.torch_dtype(rec.precision)
    Xi, _ = rec.device_tape(dev)
    x0 = _device.to_vec(s0, dt, dev)
    S_ = _device.to_colmajor(S, dt, dev)
    G_ = _device.to_colmajor(G, dt, dev)
    if tuple(S_.shape) != (D, T) or tuple(G_.shape) != (D, T):
        raise DimensionMismatch("S and G must be D x T")
    acc = torch.as_tensor(accepted).to(device=dev, dtype=torch.uint8).contiguous()
    if acc.numel() != T:
        raise DimensionMismatch("accepted must have T entries")
    dx0 = torch.empty(D, dtype=dt, device=dev)
    lam = _device.colmajor_empty(D, T, dt, dev) if return_lambda else None
    params = _lib.DeerParams(float(rec.eps), float(tol_abs), float(tol_rel), 1.0, int(maxiter),
                             _lib.JAC_STOCH_DIAG, 1, rec.precision, int(seed) & 0xFFFFFFFFFFFFFFFF)
    grad = (C.c_double * 2)()
    info = _lib.DeerInfo()
    ws.ctx.call("pmcmc_deer_adjoint", ws.h, rec.target.handle(ws.ctx), C.byref(params), _device.ptr(x0),
                _device.ptr(Xi), _device.ptr(S_), _device.ptr(acc), _device.ptr(G_), grad, _device.ptr(dx0),
                _device.ptr(lam), C.byref(info))
    inf = Info(bool(info.converged), int(info.iters), float(info.metric), "adjoint", 1.0, 1)
    out = (float(grad[0]), float(grad[1]), dx0, inf)
    return out + (lam,) if return_lambda else out


def deer_update(rec: TapedRecursion, s0, S, jacobian="stoch_diag", damping=1.0, probes=1, seed=0):
    """One Newton step from the iterate S (src/DEER/DEER.jl:477-492), with
    the probes of iteration 0 of ``seed``."""
    return solve(rec, s0, init=S, tol_abs=0.0, tol_rel=0.0, maxiter=1, jacobian=jacobian, damping=damping,
                 probes=probes, seed=seed, copy_result=True)
