"""Gradients of trajectory statistics through DEER-solved MALA.

The forward pass solves the taped MALA trajectory in windows of ``window``
steps (each window is a DEER solve started from the previous window's last
state) and keeps only each window's start state as a checkpoint.  The
backward pass walks the windows in reverse: it re-solves a window from its
checkpoint, runs the adjoint of that window (``DEER.adjoint``: reverse-time
DEER sweeps over the transposed Jacobians, accept decisions held fixed) with
the incoming dL/d(last state) added to the window's last column, and passes
dL/d(window start) on to the previous window.

``DEERTrajectory`` wraps this as a ``torch.autograd.Function`` (the role a
ChainRulesCore ``rrule`` plays in Julia): its inputs are eps, the target's
hyperparameter (the logistic prior scale) and x0, its output the D x T
trajectory.  No reference counterpart: the reference differentiates
sequential MALA with Enzyme (ext/EnzymeExt.jl).
"""
from __future__ import annotations

import numpy as np
import torch

from . import DEER, _device, _lib
from ._lib import ArgumentError


def _windows(T, window):
    return [(b, min(b + window, T)) for b in range(0, T, window)]


def _rec(model, eps, Xi, U, b, e, precision):
    return DEER.TapedRecursion(model, eps, Xi[:, b:e], U[b:e], precision=precision)


class _Workspaces:
    """One DEERWorkspace per window length, reused across windows."""

    def __init__(self):
        self.ws = {}

    def get(self, rec):
        key = (rec.D, rec.T, rec.precision)
        if key not in self.ws:
            self.ws[key] = DEER.DEERWorkspace(rec.D, rec.T, rec.precision)
        return self.ws[key]


def solve_windowed(model, eps, Xi, U, x0, window=4096, precision=None, seed=0, **solve_kw):
    """DEER-solve the trajectory window by window.  Returns (S, checkpoints,
    infos): S is D x T on the device, checkpoints[k] the start state of
    window k.  ``solve_kw`` goes to DEER.solve (tolerances, jacobian, ...)."""
    D, T = Xi.shape
    if window < 1:
        raise ArgumentError("window must be >= 1")
    xs = x0
    S_parts, ckpts, infos = [], [], []
    wss = _Workspaces()
    for k, (b, e) in enumerate(_windows(T, window)):
        rec = _rec(model, eps, Xi, U, b, e, precision)
        Sk, info = DEER.solve(rec, xs, seed=seed + k, return_info=True, workspace=wss.get(rec), **solve_kw)
        ckpts.append(_device.to_vec(xs, Sk.dtype, Sk.device).clone())
        S_parts.append(Sk)
        infos.append(info)
        xs = Sk[:, -1]
    S = torch.cat(S_parts, dim=1)
    return _device.to_colmajor(S, S.dtype, S.device), ckpts, infos


def backward_windowed(model, eps, Xi, U, ckpts, G, window=4096, precision=None, seed=0, adjoint_kw=None,
                      **solve_kw):
    """Adjoint sweep over the windows in reverse, re-solving each window from
    its checkpoint.  G = dL/dS (D x T).  Returns (dL/deps, dL/dhyper, dL/dx0,
    adjoint infos)."""
    adjoint_kw = adjoint_kw or {}
    D, T = Xi.shape
    wins = _windows(T, window)
    if len(ckpts) != len(wins):
        raise ArgumentError("one checkpoint per window expected")
    G = torch.as_tensor(G) if not isinstance(G, torch.Tensor) else G
    deps, dhyp, carry = 0.0, 0.0, None
    infos = []
    wss = _Workspaces()
    for k in range(len(wins) - 1, -1, -1):
        b, e = wins[k]
        rec = _rec(model, eps, Xi, U, b, e, precision)
        ws = wss.get(rec)
        Sk, info = DEER.solve(rec, ckpts[k], seed=seed + k, return_info=True, workspace=ws, **solve_kw)
        Gk = _device.to_colmajor(G[:, b:e], Sk.dtype, Sk.device)
        if carry is not None:
            Gk[:, -1] += carry
        de, dh, dx0, ainfo = DEER.adjoint(rec, ckpts[k], Sk, info.accepted, Gk, workspace=ws, **adjoint_kw)
        deps += de
        dhyp += dh
        carry = dx0
        infos.append(ainfo)
    return deps, dhyp, carry, infos[::-1]


class DEERTrajectory(torch.autograd.Function):
    """S = DEERTrajectory.apply(eps, hyper, x0, model, Xi, U, window, opts).

    eps and hyper are 0-d tensors (hyper must equal the logistic target's
    prior scale; pass any value for the Gaussian targets, its gradient is 0),
    x0 a length-D tensor.  opts: dict with optional keys ``precision``,
    ``seed``, ``solve`` (DEER.solve keywords) and ``adjoint`` (DEER.adjoint
    keywords).  Backward returns dL/deps, dL/dhyper and dL/dx0."""

    @staticmethod
    def forward(ctx, eps, hyper, x0, model, Xi, U, window, opts):
        opts = dict(opts or {})
        tgt = getattr(model, "target", None)
        ps = getattr(tgt, "prior_scale", None)
        if ps is not None and abs(float(hyper) - ps) > 1e-12 * max(1.0, abs(ps)):
            raise ArgumentError("hyper must equal the target's prior_scale")
        eps_f = float(eps)
        S, ckpts, infos = solve_windowed(model, eps_f, Xi, U, x0.detach(), window, opts.get("precision"),
                                         opts.get("seed", 0), **opts.get("solve", {}))
        if not all(i.converged for i in infos):
            raise _lib.PmcmcError("DEER did not converge in some window; the adjoint needs the fixed point")
        ctx.model, ctx.Xi, ctx.U, ctx.window, ctx.opts, ctx.eps = model, Xi, U, window, opts, eps_f
        ctx.ckpts = ckpts
        ctx.x0_meta = (x0.dtype, x0.device)
        return S

    @staticmethod
    def backward(ctx, gS):
        o = ctx.opts
        deps, dh, dx0, _ = backward_windowed(ctx.model, ctx.eps, ctx.Xi, ctx.U, ctx.ckpts, gS, ctx.window,
                                             o.get("precision"), o.get("seed", 0), o.get("adjoint"),
                                             **o.get("solve", {}))
        dt, dev = ctx.x0_meta
        return (torch.tensor(deps, dtype=torch.float64), torch.tensor(dh, dtype=torch.float64),
                dx0.to(dtype=dt, device=dev), None, None, None, None, None)


def trajectory_mean_sq_grad(model, eps, Xi, U, x0, window=4096, precision=None, seed=0, solve_kw=None,
                            adjoint_kw=None):
    """L = (1/T) sum_t |x_t|^2 and its gradients (dL/deps, dL/dhyper, dL/dx0)
    by the windowed checkpointed adjoint (the objective of the adjoint
    workload)."""
    solve_kw = solve_kw or {}
    S, ckpts, infos = solve_windowed(model, eps, Xi, U, x0, window, precision, seed, **solve_kw)
    T = S.shape[1]
    Sd = S.double()
    L = float((Sd * Sd).sum() / T)
    G = (2.0 / T) * S
    deps, dh, dx0, ainfos = backward_windowed(model, eps, Xi, U, ckpts, G, window, precision, seed, adjoint_kw,
                                              **solve_kw)
    return L, deps, dh, dx0, infos, ainfos
