"""Many chains of uneven D and T in one batched DEER solve.

Each chain is a taped MALA recursion (a DensityModel with a logistic or
dense-Gaussian GPU target, its own eps, x0 and tape).  Each chain's Newton
iterations run over its own sliding window of at most ``window`` columns,
which moves on as its prefix becomes final or the whole window meets the
tolerance; every iteration of every live chain runs as one set of batched
kernel launches over per-chain descriptor tables, replayed from a captured
hipGraph, and each chain stops at the end of its own trajectory
(include/pmcmc_amd.h, pmcmc_batch_*).  The reference solves one
chain per call (src/interface.jl:475-515) and has no windowing.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _device, _lib
from ._lib import ArgumentError
from .interface import _target_of


@dataclass
class Chain:
    """One chain: model (DensityModel with a GPU target), eps, x0 (D), tape
    Xi (D x T) and U (T), numpy arrays or device tensors."""

    model: object
    eps: float
    x0: object
    Xi: object
    U: object


def tune_chain_widths(chains, window, prec, seed=0):
    """Autotune the level-0 tile of the single-chain DEER solves of these
    chains (one window at a time): one representative per bucket of 16
    consecutive widths (its widest chain, on its first window), its shape
    recorded for every chain of the bucket (a shape that fits the widest chain
    fits the narrower ones: the tile's LDS grows with D).  The batched solve
    keeps the fixed rule (see batch_tile_shape in the library).  Returns
    {bucket width: shape}."""
    from . import DEER, Tune

    buckets = {}
    for ch in chains:
        D = int(np.asarray(ch.x0).shape[0])
        buckets.setdefault((D + 15) // 16 * 16, []).append(ch)
    out = {}
    for D16, group in sorted(buckets.items()):
        rep = max(group, key=lambda c: (int(np.asarray(c.x0).shape[0]), np.asarray(c.U).shape[0]))
        T = np.asarray(rep.U).shape[0]
        L = min(window, T)
        Xi = np.asarray(rep.Xi)[:, :L]
        U = np.asarray(rep.U)[:L]
        rec = DEER.TapedRecursion(rep.model, rep.eps, Xi, U, precision=prec)
        shape = Tune.tune_solve(rec, np.asarray(rep.x0), iters=3, seed=seed)
        for c in group:
            Lc = min(window, np.asarray(c.U).shape[0])
            Tune.set_entry("deer", "f32", int(np.asarray(c.x0).shape[0]), Lc, shape["tc"], shape["parts"],
                           us=shape["us"])
        out[D16] = shape
    return out


class BatchPlan:
    """Device buffers and descriptor tables for repeated solves of the same
    chains (the tapes are uploaded once).  ``solve()`` returns per-chain
    (S, accepted, Info)."""

    def __init__(self, chains, window=4096, precision="bf16_jac", tol_abs=1e-6, tol_rel=1e-5, maxiter=200,
                 damping=1.0, seed=0, use_graph=True, device=0):
        from .DEER import Info  # noqa: F401  (imported for the return type)

        if not chains:
            raise ArgumentError("no chains")
        prec = {"f32": _lib.PREC_F32, "bf16_jac": _lib.PREC_BF16_JAC}.get(str(precision).lstrip(":"))
        if prec is None:
            raise ArgumentError("batched solves run in f32 or bf16_jac")
        if window < 1 or maxiter < 1 or not (0 < damping <= 1) or tol_abs < 0 or tol_rel < 0:
            raise ArgumentError("bad window / maxiter / damping / tolerances")
        self.ctx = _device.context(device)
        dev = torch.device("cuda", device)
        self.damping, self.seed = damping, seed
        self.S, self.acc, self._keep, descs = [], [], [], []
        for ch in chains:
            tgt = _target_of(ch.model)
            Xi = _device.to_colmajor(ch.Xi, torch.float32, dev)
            D, T = Xi.shape
            if D != tgt.dim:
                raise ArgumentError("tape rows must equal the target dimension")
            U = _device.to_vec(ch.U, torch.float32, dev)
            x0 = _device.to_vec(ch.x0, torch.float32, dev)
            if U.numel() != T or x0.numel() != D:
                raise ArgumentError("U must have T entries and x0 D entries")
            S = _device.colmajor_empty(D, T, torch.float32, dev)
            acc = torch.zeros(T, dtype=torch.uint8, device=dev)
            self.S.append(S)
            self.acc.append(acc)
            self._keep += [Xi, U, x0, tgt]
            descs.append(_lib.ChainDesc(tgt.handle(self.ctx).value, T, float(ch.eps), x0.data_ptr(), Xi.data_ptr(),
                                        U.data_ptr(), S.data_ptr(), acc.data_ptr()))
        self.n = len(descs)
        arr = (_lib.ChainDesc * self.n)(*descs)
        params = _lib.DeerParams(0.0, float(tol_abs), float(tol_rel), float(damping), int(maxiter),
                                 _lib.JAC_STOCH_DIAG, 1, prec, int(seed) & 0xFFFFFFFFFFFFFFFF)
        torch.cuda.synchronize(dev)
        h = C.c_void_p()
        lib = _lib.load()
        _lib.check(lib.pmcmc_batch_create(self.ctx.h, self.n, arr, C.byref(params), int(window), C.byref(h)),
                   self.ctx.h, "pmcmc_batch_create")
        self.h = h
        _lib.check(lib.pmcmc_batch_set_graph(self.h, 1 if use_graph else 0), self.ctx.h, "pmcmc_batch_set_graph")

    def solve(self):
        from .DEER import Info

        infos = (_lib.DeerInfo * self.n)()
        torch.cuda.synchronize()
        _lib.check(_lib.load().pmcmc_batch_solve(self.h, infos), self.ctx.h, "pmcmc_batch_solve")
        return [(self.S[i], self.acc[i],
                 Info(bool(infos[i].converged), int(infos[i].iters), float(infos[i].metric), "stoch_diag",
                      self.damping, 1, self.acc[i]))
                for i in range(self.n)]

    def set_profiling(self, on=True):
        """Measure the dense-Gaussian Lambda passes and the logistic GEMM passes
        of the next solves: HIP events around every launch (direct launches,
        no graph replay) and their algorithmic flops counted on the device
        (``pass_stats``, ``glm_pass_stats``)."""
        _lib.check(_lib.load().pmcmc_batch_set_profiling(self.h, 1 if on else 0), self.ctx.h,
                   "pmcmc_batch_set_profiling")

    def pass_stats(self):
        """(launches, total_ms, flops) of the last profiled solve's Lambda passes."""
        n = C.c_int64()
        ms = C.c_double()
        fl = C.c_double()
        _lib.check(_lib.load().pmcmc_batch_pass_stats(self.h, C.byref(n), C.byref(ms), C.byref(fl)), self.ctx.h,
                   "pmcmc_batch_pass_stats")
        return int(n.value), float(ms.value), float(fl.value)

    def glm_pass_stats(self):
        """(launches, total_ms, flops) of the last profiled solve's logistic GEMM
        passes (one launch per padded width class and pass)."""
        n = C.c_int64()
        ms = C.c_double()
        fl = C.c_double()
        _lib.check(_lib.load().pmcmc_batch_glm_pass_stats(self.h, C.byref(n), C.byref(ms), C.byref(fl)),
                   self.ctx.h, "pmcmc_batch_glm_pass_stats")
        return int(n.value), float(ms.value), float(fl.value)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().pmcmc_batch_destroy(self.h)
        except Exception:
            pass


def solve_batched(chains, window=4096, precision="bf16_jac", **kw):
    """One batched solve; returns [(S, accepted, Info)] in chain order."""
    return BatchPlan(chains, window, precision, **kw).solve()
