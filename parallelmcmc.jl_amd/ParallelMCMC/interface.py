"""Sampler surface: DensityModel, ParallelMALASampler, sample, step.

Mirrors src/interface.jl (DensityModel :36-70, ParallelMALASampler :248-290,
the block driver :475-524 and :661-711, AbstractMCMC.step :811-869) with the
same names, defaults and argument checks.  Differences, all forced by
running the solve inside HIP kernels:

* ``DensityModel`` takes a GPU target descriptor (``target=``); the closure
  callbacks are accepted for signature compatibility but not called.
* The tape is drawn from a numpy Generator (Julia's MersenneTwister is not
  available); probes come from a seed drawn from the same generator once
  per block, so a block's tape does not depend on the previous block's
  iteration count (the reference threads one rng through both,
  src/interface.jl:486, 510).
* ``backend`` (the AD backend for the HVP fallback) is accepted and ignored:
  HVPs are analytic in the kernels.
* ``cholM`` (the mass matrix M = L L') is applied by reparametrization: each
  block is solved for theta = L^-1 x on the target p(L theta) (a dense
  Gaussian, or the logistic target with design X L and prior precision
  L' L / s^2), where standard MALA driven by the same tape is the
  preconditioned MALA of src/MALA/MALA.jl:29-47, 430-438 on x; the
  trajectory is mapped back by x = L theta on the device.  The fixed point
  (the trajectory) is the reference's; the Newton iterates differ (the
  Jacobian estimators act in theta coordinates).
* ``tape_rng="device"`` (opt-in, no reference counterpart) draws each
  block's tape on the GPU (``make_mala_tape_block_device``: Philox4x64-10
  from one seed drawn from the Generator) instead of on the host and
  uploading it; the default ``"host"`` keeps the numpy stream.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _device, _lib
from ._lib import ArgumentError, DimensionMismatch


class DensityModel:
    """src/interface.jl:36-70 plus the GPU target descriptor."""

    def __init__(self, logdensity=None, grad_logdensity=None, dim=None, *, param_names=None, hvp=None,
                 logdensity_batch=None, grad_logdensity_batch=None, hvp_batch=None, target=None):
        if target is not None and dim is None:
            dim = target.dim
        if dim is None:
            raise ArgumentError("dim is required")
        if target is not None and target.dim != dim:
            raise DimensionMismatch(f"target dim {target.dim} != model dim {dim}")
        self.logdensity = logdensity
        self.grad_logdensity = grad_logdensity
        self.hvp = hvp
        self.logdensity_batch = logdensity_batch
        self.grad_logdensity_batch = grad_logdensity_batch
        self.hvp_batch = hvp_batch
        self.dim = int(dim)
        self.param_names = param_names
        self.target = target


def _target_of(model):
    tgt = getattr(model, "target", None)
    if tgt is None and hasattr(model, "handle"):
        tgt = model
    if tgt is None:
        raise ArgumentError(
            "this DensityModel has no GPU target descriptor (target=...); user closures cannot run in HIP kernels"
        )
    return tgt


_PREC_NAMES = {"f64": _lib.PREC_F64, "float64": _lib.PREC_F64, "f32": _lib.PREC_F32, "float32": _lib.PREC_F32,
               "bf16_jac": _lib.PREC_BF16_JAC}


def _precision_of(eps, precision=None):
    """The reference's precision rule is FP = typeof(float(epsilon))
    (src/interface.jl:276-277): float64 unless epsilon is float32, unless a
    precision is named explicitly."""
    if precision is None:
        return _lib.PREC_F32 if isinstance(eps, (np.float32,)) else _lib.PREC_F64
    if isinstance(precision, str):
        if precision not in _PREC_NAMES:
            raise ArgumentError(f"unknown precision {precision!r}")
        return _PREC_NAMES[precision]
    return int(precision)


@dataclass
class ParallelMALASampler:
    """DEER-parallelised MALA (src/interface.jl:248-290).  Defaults follow the
    reference: T=64, maxiter=200, tol_abs=1e-6, tol_rel=1e-5,
    jacobian=:stoch_diag, damping=0.5, probes=1."""

    epsilon: float
    T: int = 64
    maxiter: int = 200
    tol_abs: float = 1e-6
    tol_rel: float = 1e-5
    jacobian: str = "stoch_diag"
    damping: float = 0.5
    probes: int = 1
    cholM: object = None
    backend: object = None
    precision: object = None
    tape_rng: str = "host"

    def __post_init__(self):
        if self.tape_rng not in ("host", "device"):
            raise ArgumentError("tape_rng must be \"host\" or \"device\"")
        if not self.epsilon > 0:
            raise ArgumentError(f"epsilon must be > 0, got {self.epsilon}")
        j = str(self.jacobian).lstrip(":")
        if j not in ("stoch_diag", "diag", "full", "full_sigmoid"):
            raise ArgumentError("jacobian must be :stoch_diag, :diag, :full or :full_sigmoid")
        self.jacobian = j
        if self.cholM is not None:
            # a Cholesky factor object (with .L, as Julia's Cholesky) or its
            # lower-triangular factor L of the mass matrix M = L L'
            L = np.asarray(getattr(self.cholM, "L", self.cholM), dtype=np.float64)
            if L.ndim != 2 or L.shape[0] != L.shape[1]:
                raise ArgumentError("cholM must be a D x D lower-triangular Cholesky factor")
            if np.any(np.triu(L, 1) != 0) or not np.all(np.diag(L) > 0):
                raise ArgumentError("cholM must be lower-triangular with a positive diagonal")
            self._L = L
            self._theta_models = {}
        self.precision = _precision_of(self.epsilon, self.precision)


@dataclass
class ParallelMALATransition:
    x: np.ndarray
    logp: float


@dataclass
class ParallelMALAState:
    x: np.ndarray
    logp: float
    trajectory: torch.Tensor
    logps: np.ndarray
    workspace: object
    tape: tuple
    t: int


class ParallelMALABlockSamples:
    """Lazy vector of transitions over kept blocks (src/interface.jl:313-332)."""

    def __init__(self, blocks, logps, n, T):
        self.blocks, self.logps, self.n, self.T = blocks, logps, n, T

    def __len__(self):
        return self.n

    def __getitem__(self, i):
        if not (0 <= i < self.n):
            raise IndexError(i)
        b, t = divmod(i, self.T)
        return ParallelMALATransition(self.blocks[b][:, t].copy(), float(self.logps[b][t]))

    def array(self):
        """N x D matrix of samples."""
        return np.concatenate([b.T for b in self.blocks], axis=0)[: self.n]


def make_mala_tape_block(rng, D, T, precision):
    """xi_t ~ N(0,I), u_t ~ U(0,1) for one block (src/interface.jl:334-358)."""
    dt = np.float64 if precision == _lib.PREC_F64 else np.float32
    Xi = np.asfortranarray(rng.standard_normal((T, D)).T.astype(dt))
    U = rng.random(T).astype(dt)
    return Xi, U


def make_mala_tape_block_device(seed, D, T, precision, device=None):
    """One block's tape drawn on the device (opt-in; the reference draws it
    on the host, src/interface.jl:334-358): Xi (D x T) ~ N(0, I) and U (T) ~
    U(0, 1) as device tensors, a pure function of ``seed``
    (pmcmc_tape_generate: Philox4x64-10, Box-Muller)."""
    dev = torch.device("cuda", device if device is not None else torch.cuda.current_device())
    dt = _device.torch_dtype(precision)
    Xi = _device.colmajor_empty(D, T, dt, dev)
    U = torch.empty(T, dtype=dt, device=dev)
    _device.context(dev.index).call("pmcmc_tape_generate", int(seed) & 0xFFFFFFFFFFFFFFFF, int(precision),
                                    _device.ptr(Xi), _device.ptr(U), int(D), int(T))
    return Xi, U


def _theta_model(model, sampler):
    """With a mass matrix M = L L' (cholM) the block is solved for
    theta = L^-1 x on the reparametrized target (targets.reparametrized):
    standard MALA there, driven by the same tape, is the preconditioned MALA
    of src/MALA/MALA.jl:29-47, 430-438, 467-471 on x.  Cached per model."""
    if sampler.cholM is None:
        return model
    key = id(model)
    if key not in sampler._theta_models:
        from .targets import reparametrized

        tgt = reparametrized(_target_of(model), sampler._L)
        sampler._theta_models[key] = (model, DensityModel(dim=model.dim, target=tgt))
    return sampler._theta_models[key][1]


def _solve_block(rng, model, sampler, x0, ws):
    from . import DEER

    D, T = model.dim, sampler.T
    if sampler.cholM is not None:
        x0 = np.linalg.solve(sampler._L, np.asarray(x0, dtype=np.float64))
        model = _theta_model(model, sampler)
    if sampler.tape_rng == "device":
        Xi, U = make_mala_tape_block_device(int(rng.integers(0, 2**63 - 1)), D, T, sampler.precision)
    else:
        Xi, U = make_mala_tape_block(rng, D, T, sampler.precision)
    seed = int(rng.integers(0, 2**63 - 1))
    rec = DEER.TapedRecursion(model, sampler.epsilon, Xi, U, sampler.precision)
    if ws is None:
        ws = DEER.DEERWorkspace(D, T, sampler.precision)
    S, info = DEER.solve(rec, x0, tol_abs=sampler.tol_abs, tol_rel=sampler.tol_rel, maxiter=sampler.maxiter,
                         jacobian=sampler.jacobian, damping=sampler.damping, probes=sampler.probes, seed=seed,
                         return_info=True, workspace=ws, copy_result=False)
    return S, (Xi, U), ws, info


def _trajectory_logps(model, S, precision):
    from . import MALA

    return MALA.logdensity_batch(model, S, precision=precision).cpu().numpy()


def _block_trajectory(model, S, ws, sampler):
    """The block's trajectory and log-densities (src/interface.jl:517-524).
    GEMM targets (logistic, dense Gaussian): both from the final Newton
    iteration, assembled on the device (DEER.last_trajectory: FT of the last
    iterate, within the solve's tolerance of S, with the log-densities its
    passes already evaluated).  Diagonal targets: S and one elementwise
    log-density kernel."""
    from . import DEER
    from .owned import pmcmc_matmul
    from .targets import GaussianTarget, LogisticTarget

    m = _theta_model(model, sampler)
    if isinstance(m.target, (LogisticTarget, GaussianTarget)):
        S, lp = DEER.last_trajectory(ws)
        lp = lp.cpu().numpy()
    else:
        lp = _trajectory_logps(m, S, sampler.precision)
    if sampler.cholM is not None:
        # x = L theta (device GEMM); the log-densities are those of x already
        L = torch.as_tensor(sampler._L, dtype=S.dtype, device=S.device)
        S = pmcmc_matmul(L, S)
    return S, lp


def sample(rng, model, sampler, N, initial_params=None, progress=False, return_infos=False):
    """Draw N samples in blocks of sampler.T (src/interface.jl:661-711)."""
    if N < 1:
        raise ArgumentError("the number of samples must be >= 1")
    if not isinstance(rng, np.random.Generator):
        rng = np.random.default_rng(rng)
    D = model.dim
    x0 = np.asarray(initial_params, dtype=np.float64).copy() if initial_params is not None else rng.standard_normal(D)
    blocks, logp_blocks, infos = [], [], []
    ws = None
    nsteps = 0
    while nsteps < N:
        S, tape, ws, info = _solve_block(rng, model, sampler, x0, ws)
        nkeep = min(sampler.T, N - nsteps)
        S, logps = _block_trajectory(model, S, ws, sampler)
        S_keep = S[:, :nkeep].cpu().numpy().astype(np.float64, copy=True)
        logps = logps[:nkeep]
        blocks.append(S_keep)
        logp_blocks.append(logps)
        infos.append(info)
        x0 = S_keep[:, nkeep - 1].copy()
        nsteps += nkeep
    out = ParallelMALABlockSamples(blocks, logp_blocks, N, sampler.T)
    return (out, infos) if return_infos else out


def step(rng, model, sampler, state=None, initial_params=None):
    """AbstractMCMC.step (src/interface.jl:811-869): walk the cached
    trajectory, re-solving from its last column when it is exhausted."""
    if not isinstance(rng, np.random.Generator):
        rng = np.random.default_rng(rng)
    if state is None or state.t + 1 >= sampler.T:
        if state is None:
            x0 = (np.asarray(initial_params, dtype=np.float64).copy() if initial_params is not None
                  else rng.standard_normal(model.dim))
            ws = None
        else:
            x0 = state.trajectory[:, sampler.T - 1].cpu().numpy().astype(np.float64)
            ws = state.workspace
        S, tape, ws, _ = _solve_block(rng, model, sampler, x0, ws)
        S, logps = _block_trajectory(model, S, ws, sampler)
        S = S.t().clone().t()
        x = S[:, 0].cpu().numpy().astype(np.float64)
        st = ParallelMALAState(x, float(logps[0]), S, logps, ws, tape, 0)
        return ParallelMALATransition(x, float(logps[0])), st
    t = state.t + 1
    x = state.trajectory[:, t].cpu().numpy().astype(np.float64)
    st = ParallelMALAState(x, float(state.logps[t]), state.trajectory, state.logps, state.workspace, state.tape, t)
    return ParallelMALATransition(x, float(state.logps[t])), st
