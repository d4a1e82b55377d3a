"""CPU oracle for the DEER-over-MALA path (numpy, FP64 unless a dtype is given).

TEST INFRASTRUCTURE ONLY.  This module restates, in plain numpy, the reference
algorithm of rsenne/ParallelMCMC.jl for the one path this repository accelerates
(DEER's Newton solve over taped MALA trajectories).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or as the timed CPU baseline.  The product
path (``parallelmcmc.jl_amd``) never imports it and has no CPU fallback.

Every function cites the reference file:line it restates.  How each one is
pinned (closed forms, invariants, brute force, finite differences) is listed in
``PINNED`` at the bottom and exercised by ``tests/test_oracle_*.py``.

Layout follows the reference: a trajectory is a ``D x T`` matrix whose column
``t`` is the state after step ``t``; ``s0`` is the state before step 1.

Deliberate, documented deviations from the reference:

* Rademacher probes.  The reference draws them from the sampler's host RNG
  (src/DEER/DEER.jl:298-325, 407, 412).  Here a probe entry is a pure function
  of ``(seed, iteration, probe, t*D + d)`` (``rademacher``), so the GPU kernels
  can generate the identical matrix on chip.  Probes only change the Newton
  iteration count, never the fixed point (src/DEER/DEER.jl:409-419).
* Logistic log-likelihood uses the stable softplus form; it is the same
  function as benchmarks/.../bayes_logreg.jl:130-131.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ----------------------------------------------------------------------------
# Targets.  The reference's DensityModel holds user closures
# (src/interface.jl:36-70); the hot path needs logdensity/grad/hvp and their
# batched forms.  These are the analytic targets the GPU kernels implement.
# ----------------------------------------------------------------------------


def _softplus(x):
    return np.maximum(x, 0.0) + np.log1p(np.exp(-np.abs(x)))


def _sigmoid(x):
    out = np.empty_like(x)
    pos = x >= 0
    out[pos] = 1.0 / (1.0 + np.exp(-x[pos]))
    ex = np.exp(x[~pos])
    out[~pos] = ex / (1.0 + ex)
    return out


class Target:
    """Common batched interface: X is D x M (columns are states)."""

    dim: int

    def logp(self, x):
        return float(self.logp_batch(x.reshape(-1, 1))[0])

    def grad(self, x):
        return self.grad_batch(x.reshape(-1, 1))[:, 0]

    def hvp(self, x, v):
        return self.hvp_batch(x.reshape(-1, 1), v.reshape(-1, 1))[:, 0]


@dataclass
class DiagGaussian(Target):
    """N(mean, diag(1/prec)): logp = -1/2 sum prec (x-mean)^2 (constant dropped).

    The standard normal of test/test-MALA-Kernel.jl:9-10 and
    test/test-Deer-vs-MALA.jl:13-14 is prec=1, mean=0.
    """

    prec: np.ndarray
    mean: np.ndarray

    @property
    def dim(self):
        return len(self.prec)

    @staticmethod
    def std_normal(D, dtype=np.float64):
        return DiagGaussian(np.ones(D, dtype), np.zeros(D, dtype))

    def logp_batch(self, X):
        R = X - self.mean[:, None]
        return -0.5 * np.sum(self.prec[:, None] * R * R, axis=0)

    def grad_batch(self, X):
        return -self.prec[:, None] * (X - self.mean[:, None])

    def hvp_batch(self, X, V):
        return -self.prec[:, None] * V

    def hess_diag_batch(self, X):
        return np.broadcast_to(-self.prec[:, None], X.shape).copy()


@dataclass
class Gaussian(Target):
    """N(mean, Lam^-1) with a dense symmetric precision Lam."""

    Lam: np.ndarray
    mean: np.ndarray

    @property
    def dim(self):
        return self.Lam.shape[0]

    def logp_batch(self, X):
        R = X - self.mean[:, None]
        return -0.5 * np.sum(R * (self.Lam @ R), axis=0)

    def grad_batch(self, X):
        return -(self.Lam @ (X - self.mean[:, None]))

    def hvp_batch(self, X, V):
        return -(self.Lam @ V)

    def hess_diag_batch(self, X):
        return np.broadcast_to(-np.diag(self.Lam)[:, None], X.shape).copy()


@dataclass
class LogReg(Target):
    """Bayesian logistic regression, prior N(0, prior_scale^2 I).

    Restates benchmarks/ParallelMCMCBenchmarks/src/models/bayes_logreg.jl:118-203
    (logp_batch 118-137, gradlogp_batch 139-154, hvp_batch 181-203); the
    reference's prior is prior_scale = 1.
    """

    Xd: np.ndarray  # N x D design
    y: np.ndarray  # N labels in {0,1}
    prior_scale: float = 1.0

    @property
    def dim(self):
        return self.Xd.shape[1]

    def logp_batch(self, B):
        L = self.Xd @ B
        y = self.y[:, None]
        ll = -y * _softplus(-L) - (1.0 - y) * _softplus(L)
        return ll.sum(axis=0) - 0.5 * np.sum(B * B, axis=0) / self.prior_scale**2

    def grad_batch(self, B):
        P = _sigmoid(self.Xd @ B)
        return self.Xd.T @ (self.y[:, None] - P) - B / self.prior_scale**2

    def hvp_batch(self, B, V):
        P = _sigmoid(self.Xd @ B)
        W = P * (1.0 - P)
        return -(self.Xd.T @ (W * (self.Xd @ V))) - V / self.prior_scale**2

    def hess_diag_batch(self, B):
        P = _sigmoid(self.Xd @ B)
        W = P * (1.0 - P)
        return -((self.Xd * self.Xd).T @ W) - 1.0 / self.prior_scale**2

    def dgrad_dhyper(self, B):
        """d grad / d prior_scale = 2 B / prior_scale^3 (the prior's term)."""
        return 2.0 * B / self.prior_scale**3


def make_logreg_data(rng, N, D, dtype=np.float64):
    """Synthetic logistic data: X_ij ~ N(0,1)/sqrt(D), theta* ~ N(0,I),
    y ~ Bernoulli(sigmoid(X theta*)).  Mirrors bayes_logreg.jl:17-23 with the
    1/sqrt(D) design scale of the logistic workload."""
    theta = rng.standard_normal(D)
    X = rng.standard_normal((N, D)) / math.sqrt(D)
    p = 1.0 / (1.0 + np.exp(-(X @ theta)))
    y = (rng.random(N) < p).astype(np.float64)
    return X.astype(dtype), y.astype(dtype), theta


def make_corr_gaussian(rng, D, cond_log10=3.0):
    """Precision Lam = Q diag(10^U(0,cond)) Q^T with Q Haar-orthogonal."""
    G = rng.standard_normal((D, D))
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))[None, :]
    ev = 10.0 ** rng.uniform(0.0, cond_log10, D)
    Lam = (Q * ev[None, :]) @ Q.T
    return 0.5 * (Lam + Lam.T), ev


# ----------------------------------------------------------------------------
# Sequential (scalar) MALA — the oracle proper.
# ----------------------------------------------------------------------------


def logq_mala(y, x, gx, eps):
    """log q(y|x), proposal N(x + eps*g, 2 eps I).  src/MALA/MALA.jl:172-193."""
    d = len(x)
    r = y - (x + eps * gx)
    return -0.5 * np.dot(r, r) / (2 * eps) - (d / 2) * math.log(4 * math.pi * eps)


def mala_proposal(target, x, eps, xi):
    """y = x + eps*grad(x) + sqrt(2 eps) xi.  src/MALA/MALA.jl:207-229."""
    gx = target.grad(x)
    return x + eps * gx + math.sqrt(2 * eps) * xi, gx


def mala_logalpha(target, x, y, eps):
    """src/MALA/MALA.jl:245-261."""
    gx = target.grad(x)
    gy = target.grad(y)
    return (target.logp(y) + logq_mala(x, y, gy, eps)) - (target.logp(x) + logq_mala(y, x, gx, eps))


def mala_step_taped(target, x, eps, xi, u, check_u=True):
    """One taped MALA step; returns (x_next, accepted, logalpha).

    src/MALA/MALA.jl:276-305 (accept iff log(u) < logalpha, line 299) and the
    logalpha-returning form :377-411.  The scalar reference requires 0<u<1
    (line 288); the batched path does not (:523-526)."""
    if len(x) != len(xi):
        raise ValueError("x and xi must have the same length")
    if check_u and not (0.0 < u < 1.0):
        raise ValueError("u must be in (0, 1)")
    y, gx = mala_proposal(target, x, eps, xi)
    lpx = target.logp(x)
    lpy = target.logp(y)
    gy = target.grad(y)
    lq_yx = logq_mala(y, x, gx, eps)
    lq_xy = logq_mala(x, y, gy, eps)
    la = (lpy + lq_xy) - (lpx + lq_yx)
    acc = math.log(u) < la if u > 0 else True
    return (y.copy() if acc else x.copy()), acc, la


def run_mala_sequential_taped(target, x0, eps, Xi, U, check_u=True):
    """Sequential taped MALA (the oracle).  src/MALA/MALA.jl:322-346.

    Returns S (D x T, column t = x_t), accept mask (T,), logalpha (T,)."""
    D, T = Xi.shape
    if len(U) != T:
        raise ValueError("xi and u must have the same length")
    S = np.empty((D, T), dtype=np.float64)
    acc = np.zeros(T, dtype=bool)
    la = np.zeros(T)
    x = np.asarray(x0, dtype=np.float64).copy()
    for t in range(T):
        x, acc[t], la[t] = mala_step_taped(target, x, eps, Xi[:, t].astype(np.float64), float(U[t]), check_u)
        S[:, t] = x
    return S, acc, la


def logq_mala_mass(y, x, gx, eps, L):
    """log q(y|x) with the mass matrix M = L L': proposal N(x + eps M g,
    2 eps M).  src/MALA/MALA.jl:172-193 (cholM branch, _quad_Minv! :41-44,
    _logdet_M :47)."""
    d = len(x)
    M = L @ L.T
    r = y - (x + eps * (M @ gx))
    w = np.linalg.solve(L, r)
    logdet = 2.0 * np.sum(np.log(np.diag(L)))
    return -0.5 * np.dot(w, w) / (2 * eps) - (d / 2) * math.log(4 * math.pi * eps) - 0.5 * logdet


def mala_step_taped_mass(target, x, eps, xi, u, L):
    """One taped MALA step with the mass matrix M = L L':
    y = x + eps M grad(x) + sqrt(2 eps) L xi (src/MALA/MALA.jl:207-229,
    _apply_M! :31-35, _apply_L! :38), log q as logq_mala_mass, accept iff
    log u < log alpha (:276-305).  Returns (x_next, accepted, logalpha)."""
    M = L @ L.T
    gx = target.grad(x)
    y = x + eps * (M @ gx) + math.sqrt(2 * eps) * (L @ xi)
    gy = target.grad(y)
    la = (target.logp(y) + logq_mala_mass(x, y, gy, eps, L)) - (target.logp(x) + logq_mala_mass(y, x, gx, eps, L))
    acc = math.log(u) < la if u > 0 else True
    return (y.copy() if acc else x.copy()), acc, la


def run_mala_sequential_taped_mass(target, x0, eps, Xi, U, L):
    """Sequential taped MALA with the mass matrix M = L L' (src/MALA/MALA.jl:
    322-346 with cholM).  Returns S (D x T), accept mask, logalpha."""
    D, T = Xi.shape
    S = np.empty((D, T))
    acc = np.zeros(T, dtype=bool)
    la = np.zeros(T)
    x = np.asarray(x0, dtype=np.float64).copy()
    for t in range(T):
        x, acc[t], la[t] = mala_step_taped_mass(target, x, eps, Xi[:, t].astype(np.float64), float(U[t]), L)
        S[:, t] = x
    return S, acc, la


def mala_sampler_run(target, x0, eps, n, rng):
    """``n`` steps of the sequential MALASampler (the CPU baseline).

    src/interface.jl:163-194: per step a fresh xi ~ N(0, I) and u ~ U(0,1)
    from the sampler's RNG (:174-175), then mala_step_with_logalpha!
    (src/MALA/MALA.jl:377-411: proposal with grad(x), logp(x), logp(y),
    grad(y), the two log-q terms, accept iff log(u) < logalpha), and
    logdensity(x_next) re-evaluated when the step is accepted (:190).
    Returns the D x n states, the accept mask and the logp of every state."""
    x = np.asarray(x0, dtype=np.float64).copy()
    D = len(x)
    S = np.empty((D, n))
    acc = np.zeros(n, dtype=bool)
    lps = np.empty(n)
    lp = target.logp(x)
    for t in range(n):
        xi = rng.standard_normal(D)
        u = rng.random()
        y, gx = mala_proposal(target, x, eps, xi)
        lpx = target.logp(x)
        lpy = target.logp(y)
        gy = target.grad(y)
        la = (lpy + logq_mala(x, y, gy, eps)) - (lpx + logq_mala(y, x, gx, eps))
        if math.log(u) < la:
            x = y
            acc[t] = True
            lp = target.logp(x)
        S[:, t] = x
        lps[t] = lp
    return S, acc, lps


def mala_step_taped_and_jvp(target, x, eps, xi, u, v):
    """Fused exact step + JVP of the sigmoid surrogate.  src/MALA/MALA.jl:732-797
    (the JVP alone is :671-725 and returns the same jvp vector)."""
    y, gx = mala_proposal(target, x, eps, xi)
    lpx, lpy = target.logp(x), target.logp(y)
    gy = target.grad(y)
    la = (lpy + logq_mala(x, y, gy, eps)) - (lpx + logq_mala(y, x, gx, eps))
    lu = math.log(u) if u > 0 else -math.inf
    acc = lu < la
    x_next = y.copy() if acc else x.copy()
    g = 1.0 / (1.0 + math.exp(-(la - lu))) if la - lu > -700 else 0.0
    a = 1.0 if acc else 0.0
    w = v + eps * target.hvp(x, v)
    hy = target.hvp(y, w)
    r = x - y - eps * gy
    dr = v - w - eps * hy
    dla = np.sum(gy * w - gx * v - (1.0 / (2 * eps)) * r * dr)
    dg = g * (1 - g) * dla
    jvp = a * w + (y - x) * dg + (1 - a) * v
    return x_next, jvp


def mala_step_surrogate_sigmoid_jvp(target, x, eps, xi, u, v):
    """src/MALA/MALA.jl:671-725."""
    return mala_step_taped_and_jvp(target, x, eps, xi, u, v)[1]


def mala_step_surrogate_sigmoid(target, x, eps, xi, u):
    """Stop-gradient surrogate g*y + (1-g)*x.  src/MALA/MALA.jl:416-427.
    Its derivative with g's dependence on x and y kept, but a and the
    surrogate weights frozen as in :795, is what the JVP returns; used only
    by the finite-difference pins."""
    y, _ = mala_proposal(target, x, eps, xi)
    la = mala_logalpha(target, x, y, eps)
    g = 1.0 / (1.0 + math.exp(-(la - math.log(u))))
    return g * y + (1 - g) * x


# ----------------------------------------------------------------------------
# Batched MALA (columns are independent states).
# ----------------------------------------------------------------------------


def logq_mala_batched(Y, X, GX, eps):
    """src/MALA/MALA.jl:446-479 (cholM = nothing)."""
    D = X.shape[0]
    R = Y - X - eps * GX
    c = (D / 2) * math.log(4 * math.pi * eps)
    return -0.5 * np.sum(R * R, axis=0) / (2 * eps) - c


def mala_step_batched(target, X, eps, Xi, U):
    """src/MALA/MALA.jl:487-528.  Returns (X_next, accepted)."""
    GX = target.grad_batch(X)
    Y = X + eps * GX + math.sqrt(2 * eps) * Xi
    lpx, lpy = target.logp_batch(X), target.logp_batch(Y)
    GY = target.grad_batch(Y)
    la = (lpy + logq_mala_batched(X, Y, GY, eps)) - (lpx + logq_mala_batched(Y, X, GX, eps))
    with np.errstate(divide="ignore"):
        lu = np.log(U)
    acc = lu < la
    return np.where(acc[None, :], Y, X), acc


@dataclass
class FwdJvp:
    FT: np.ndarray
    Jt: np.ndarray
    acc: np.ndarray
    logalpha: np.ndarray
    Y: np.ndarray
    g: np.ndarray
    dlogalpha: np.ndarray


def mala_step_batched_fwd_and_jvp(target, X, eps, Xi, U, Z):
    """Batched exact step + surrogate JVP along Z, per column.

    src/MALA/MALA.jl:556-628 (cholM = nothing):
      Y = X + eps G_X + sqrt(2eps) Xi ; logalpha ; g = sigmoid(logalpha - log u)
      acc = log u < logalpha ; FT = acc ? Y : X
      W = Z + eps H(X) Z ; R = X - Y - eps G_Y ; dR = Z - W - eps H(Y) W
      dlogalpha = sum(G_Y W) - sum(G_X Z) - sum(R dR)/(2 eps)
      Jt = (acc ? W : Z) + (Y - X) g (1-g) dlogalpha
    """
    D, T = X.shape
    for name, M in (("Xi", Xi), ("Z", Z)):
        if M.shape != (D, T):
            raise ValueError(f"X and {name} must have the same size")
    if len(U) != T:
        raise ValueError("u must have length size(X,2)")
    GX = target.grad_batch(X)
    Y = X + eps * GX + math.sqrt(2 * eps) * Xi
    lpx, lpy = target.logp_batch(X), target.logp_batch(Y)
    GY = target.grad_batch(Y)
    la = (lpy + logq_mala_batched(X, Y, GY, eps)) - (lpx + logq_mala_batched(Y, X, GX, eps))
    with np.errstate(divide="ignore", over="ignore"):
        lu = np.log(U.astype(np.float64))
        g = 1.0 / (1.0 + np.exp(-(la - lu)))
    acc = lu < la
    FT = np.where(acc[None, :], Y, X)
    W = Z + eps * target.hvp_batch(X, Z)
    HY = target.hvp_batch(Y, W)
    R = X - Y - eps * GY
    dR = Z - W - eps * HY
    dla = np.sum(GY * W, axis=0) - np.sum(GX * Z, axis=0) - (1.0 / (2 * eps)) * np.sum(R * dR, axis=0)
    dg = g * (1 - g) * dla
    Jt = np.where(acc[None, :], W, Z) + (Y - X) * dg[None, :]
    return FwdJvp(FT, Jt, acc, la, Y, g, dla)


def mala_jac_diag_batched(target, X, eps, Xi, U):
    """Exact diag of the surrogate Jacobian J = a(I + eps H_x) + (1-a) I
    + g(1-g)(y-x) c^T, where c = grad_x logalpha of the frozen-accept
    surrogate (the map whose JVP is src/MALA/MALA.jl:749-795):
      c = (I+eps H_x) g_y - g_x - [r - (I+eps H_x)(I+eps H_y) r]/(2 eps),
      r = x - y - eps g_y.
    Equals jac_diag_via_jvps (src/DEER/DEER.jl:327-342) in exact arithmetic;
    pinned against it in the tests.  Needs three HVPs per column instead of D.
    """
    GX = target.grad_batch(X)
    Y = X + eps * GX + math.sqrt(2 * eps) * Xi
    lpx, lpy = target.logp_batch(X), target.logp_batch(Y)
    GY = target.grad_batch(Y)
    la = (lpy + logq_mala_batched(X, Y, GY, eps)) - (lpx + logq_mala_batched(Y, X, GX, eps))
    with np.errstate(divide="ignore", over="ignore"):
        lu = np.log(U.astype(np.float64))
        g = 1.0 / (1.0 + np.exp(-(la - lu)))
    acc = lu < la
    FT = np.where(acc[None, :], Y, X)
    R = X - Y - eps * GY
    HyR = target.hvp_batch(Y, R)
    Q = R + eps * HyR  # (I + eps H_y) r
    PQ = Q + eps * target.hvp_batch(X, Q)  # (I+eps H_x)(I+eps H_y) r
    PG = GY + eps * target.hvp_batch(X, GY)
    c = PG - GX - (R - PQ) / (2 * eps)
    hd = target.hess_diag_batch(X)
    a = acc.astype(np.float64)[None, :]
    Jd = a * (1 + eps * hd) + (1 - a) + (g * (1 - g))[None, :] * (Y - X) * c
    return FT, Jd, acc, la


def mala_jac_full_batched(target, X, eps, Xi, U, sigmoid=True):
    """Dense Jacobian per column.  With ``sigmoid`` the surrogate's
    J_t = a(I + eps H_x) + (1-a) I + g(1-g)(y-x) c^T with c as in
    mala_jac_diag_batched (the matrix whose columns are the JVPs of
    src/MALA/MALA.jl:749-795 along basis vectors); without it the derivative
    of the taped step with the accept decision held fixed,
    J_t = a(I + eps H_x) + (1-a) I (exact wherever no decision flips).
    No reference counterpart: the reference rejects jacobian=:full
    (src/DEER/DEER.jl:464-465).  Returns FT, J (T x D x D), acc, logalpha."""
    D, T = X.shape
    GX = target.grad_batch(X)
    Y = X + eps * GX + math.sqrt(2 * eps) * Xi
    lpx, lpy = target.logp_batch(X), target.logp_batch(Y)
    GY = target.grad_batch(Y)
    la = (lpy + logq_mala_batched(X, Y, GY, eps)) - (lpx + logq_mala_batched(Y, X, GX, eps))
    with np.errstate(divide="ignore", over="ignore"):
        lu = np.log(U.astype(np.float64))
        g = 1.0 / (1.0 + np.exp(-(la - lu)))
    acc = lu < la
    FT = np.where(acc[None, :], Y, X)
    R = X - Y - eps * GY
    Q = R + eps * target.hvp_batch(Y, R)
    PQ = Q + eps * target.hvp_batch(X, Q)
    PG = GY + eps * target.hvp_batch(X, GY)
    Cm = PG - GX - (R - PQ) / (2 * eps)
    Um = (g * (1 - g))[None, :] * (Y - X)
    if not sigmoid:
        Um = np.zeros_like(Um)
    J = np.empty((T, D, D))
    I = np.eye(D)
    for t in range(T):
        Hx = np.stack([target.hvp(X[:, t], I[:, i]) for i in range(D)], axis=1)
        a = 1.0 if acc[t] else 0.0
        J[t] = I + a * eps * Hx + np.outer(Um[:, t], Cm[:, t])
    return FT, J, acc, la


# ----------------------------------------------------------------------------
# Rademacher probes (counter based; see module docstring).
# ----------------------------------------------------------------------------

_M32 = np.uint64(0xFFFFFFFF)


def fmix32(h):
    """MurmurHash3 32-bit finaliser, on uint32 numpy arrays or ints."""
    h = np.asarray(h, dtype=np.uint64) & _M32
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & _M32
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & _M32
    h ^= h >> np.uint64(16)
    return h


def probe_key(seed, it, k):
    """Key of probe matrix k at Newton iteration it (0-based)."""
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    lo, hi = s & 0xFFFFFFFF, s >> 32
    a = int(fmix32(lo ^ 0x3C6EF372))
    b = (a + hi * 0x9E3779B1 + int(it) * 0x85EBCA77 + int(k) * 0xC2B2AE3D) & 0xFFFFFFFF
    return int(fmix32(b))


def rademacher(seed, it, k, D, T, t0=0):
    """D x T matrix of +-1; entry (d, t) is a function of (seed, it, k, (t0+t)*D+d).

    Replaces the host draw of src/DEER/DEER.jl:298-325 (z_i = rand(Bool) ? 1 : -1)."""
    key = np.uint64(probe_key(seed, it, k))
    idx = (np.arange(t0 * D, (t0 + T) * D, dtype=np.uint64)) & _M32
    h = fmix32(((idx * np.uint64(0x9E3779B1)) & _M32) ^ key)
    z = np.where((h >> np.uint64(31)) != 0, 1.0, -1.0)
    return z.reshape(T, D).T.copy()


# ----------------------------------------------------------------------------
# Affine scans.  src/DEER/DEERScan.jl.
# ----------------------------------------------------------------------------


def _check_scan_shapes(A, B, s0, S=None):
    D, T = A.shape
    if B.shape != (D, T):
        raise ValueError("B must have the same size as A")
    if S is not None and S.shape != (D, T):
        raise ValueError("S must have size (D, T)")
    if len(s0) != D:
        raise ValueError("length(s0) must equal size(A,1)")


def solve_affine_seq(A, B, s0):
    """s_t = A[:,t]*s_{t-1} + B[:,t].  src/DEER/DEERScan.jl:51-68."""
    _check_scan_shapes(A, B, s0)
    D, T = A.shape
    S = np.empty((D, T), dtype=np.result_type(A, B, s0))
    s = np.asarray(s0)
    for t in range(T):
        s = A[:, t] * s + B[:, t]
        S[:, t] = s
    return S


def solve_affine_scan_diag(A, B, s0):
    """Hillis-Steele doubling scan over affine maps,
    (a2,b2)o(a1,b1) = (a2 a1, a2 b1 + b2).  src/DEER/DEERScan.jl:102-173."""
    _check_scan_shapes(A, B, s0)
    D, T = A.shape
    if T == 0:
        return np.empty((D, 0), dtype=np.result_type(A, B, s0))
    al, be = A.copy(), B.copy()
    off = 1
    while off < T:
        an, bn = al.copy(), be.copy()
        an[:, off:] = al[:, off:] * al[:, : T - off]
        bn[:, off:] = al[:, off:] * be[:, : T - off] + be[:, off:]
        al, be = an, bn
        off <<= 1
    return al * s0[:, None] + be


def affine_scan_residual(S, A, B, s0):
    """max |S[:,t] - (A[:,t] S[:,t-1] + B[:,t])|.  src/DEER/DEERScan.jl:203-219."""
    _check_scan_shapes(A, B, s0, S)
    D, T = A.shape
    if T == 0:
        return 0.0
    prev = np.concatenate([np.asarray(s0)[:, None], S[:, :-1]], axis=1)
    return float(np.max(np.abs(S - (A * prev + B))))


def check_affine_scan(A, B, s0, atol=1e-6, rtol=1e-6, scan=None):
    """src/DEER/DEERScan.jl:236-258.  ``scan`` defaults to the doubling scan."""
    S_seq = solve_affine_seq(A, B, s0)
    S_scan = (scan or solve_affine_scan_diag)(A, B, s0)
    diff = np.abs(S_seq - S_scan)
    scale = atol + rtol * np.abs(S_seq)
    return dict(
        ok=bool(np.all(diff <= scale)),
        max_abs_err=float(diff.max()) if diff.size else 0.0,
        max_rel_err=float((diff / scale).max()) if diff.size else 0.0,
        residual_seq=affine_scan_residual(S_seq, A, B, s0),
        residual_scan=affine_scan_residual(S_scan, A, B, s0),
        S_seq=S_seq,
        S_scan=S_scan,
    )


def solve_affine_seq_full(J, B, s0):
    """Full-matrix recurrence s_t = J_t s_{t-1} + b_t, J is T x D x D.
    No reference counterpart (the reference rejects jacobian=:full,
    src/DEER/DEER.jl:464-465); this is its sequential ground truth."""
    D, T = B.shape
    S = np.empty((D, T))
    s = np.asarray(s0, dtype=np.float64)
    for t in range(T):
        s = J[t] @ s + B[:, t]
        S[:, t] = s
    return S


# ----------------------------------------------------------------------------
# Jacobian estimators.  src/DEER/DEER.jl:327-367.
# ----------------------------------------------------------------------------


def jac_diag_via_jvps(jvp, x, t):
    """Exact diag(J_t) from D JVPs along basis vectors.  src/DEER/DEER.jl:327-342."""
    D = len(x)
    d = np.zeros(D)
    for i in range(D):
        e = np.zeros(D)
        e[i] = 1.0
        d += e * jvp(x, t, e)
    return d


def jac_full_via_jvps(jvp, x, t):
    """Dense J_t (columns are JVPs along basis vectors); brute force for :full."""
    D = len(x)
    J = np.empty((D, D))
    for i in range(D):
        e = np.zeros(D)
        e[i] = 1.0
        J[:, i] = jvp(x, t, e)
    return J


def jac_diag_stoch(jvp, x, t, probes=1, rng=None, zs=None):
    """Hutchinson diag estimate mean_k z_k * (J z_k).  src/DEER/DEER.jl:344-367."""
    if probes < 1:
        raise ValueError("probes must be >= 1")
    D = len(x)
    d = np.zeros(D)
    for k in range(probes):
        z = zs[k] if zs is not None else np.where(rng.random(D) < 0.5, -1.0, 1.0)
        d += z * jvp(x, t, z)
    return d / probes


# ----------------------------------------------------------------------------
# DEER.  src/DEER/DEER.jl:375-595.
# ----------------------------------------------------------------------------


@dataclass
class DEERInfo:
    converged: bool
    iters: int
    metric: float
    jacobian: str
    damping: float
    probes: int
    accepted: np.ndarray = field(default=None, repr=False)
    logalpha: np.ndarray = field(default=None, repr=False)
    metrics: list = field(default_factory=list, repr=False)


def _xbar(s0, S):
    """Xbar = [s0, S[:, 1:T-1]].  src/DEER/DEER.jl:402-405."""
    return np.concatenate([np.asarray(s0)[:, None], S[:, :-1]], axis=1)


def deer_update_rec(fwd_and_jvp_batch, s0, S_in, jacobian="stoch_diag", damping=1.0, probes=1,
                    seed=0, it=0, scan=None, jac_diag_batch=None, jac_full_batch=None):
    """One DEER Newton step for a generic taped recursion.

    src/DEER/DEER.jl:375-475, batched branch :398-419:
      Xbar = [s0, S_in[:, 1:T-1]] ; (FT, Jt) = fwd_and_jvp_batch(Xbar, Z)
      A = mean_k Z_k * Jt_k ; B = FT - A * Xbar ; S_out = scan(A, B, s0)
      S_out = (1-damping) S_in + damping S_out
    ``fwd_and_jvp_batch(Xbar, Z) -> (FT, Jt, acc, logalpha)``.
    jacobian='diag' needs ``jac_diag_batch(Xbar) -> (FT, diagJ, acc, logalpha)``
    (the reference's scalar branch :456-462 computes the same diagonal with D
    JVPs per step).  jacobian='full' / 'full_sigmoid' (no reference
    counterpart) need ``jac_full_batch(Xbar, sigmoid) -> (FT, J, acc,
    logalpha)`` (decision-frozen / surrogate dense Jacobian) and solve the
    dense recurrence S_t = J_t S_{t-1} + FT_t - J_t Xbar_t sequentially.
    Returns (S_out, accepted, logalpha)."""
    D, T = S_in.shape
    if len(s0) != D:
        raise ValueError("S must be D x T with D = length(s0)")
    if not (0 < damping <= 1):
        raise ValueError("damping must be in (0,1]")
    if probes < 1:
        raise ValueError("probes must be >= 1")
    Xbar = _xbar(s0, S_in)
    if jacobian == "stoch_diag":
        A = np.zeros((D, T))
        for k in range(probes):
            Z = rademacher(seed, it, k, D, T)
            FT_k, Jt, acc_k, la_k = fwd_and_jvp_batch(Xbar, Z)
            A += Z * Jt
            if k == 0:
                FT, acc, la = FT_k, acc_k, la_k
        if probes > 1:
            A *= 1.0 / probes
    elif jacobian == "diag":
        FT, A, acc, la = jac_diag_batch(Xbar)
    elif jacobian in ("full", "full_sigmoid") and jac_full_batch is not None:
        FT, J, acc, la = jac_full_batch(Xbar, jacobian == "full_sigmoid")
        B = FT - np.einsum("tij,jt->it", J, Xbar)
        S_out = solve_affine_seq_full(J, B, np.asarray(s0, dtype=np.float64))
        if damping != 1.0:
            S_out = (1 - damping) * S_in + damping * S_out
        return S_out, acc, la
    else:
        raise ValueError("jacobian must be :diag or :stoch_diag")
    B = FT - A * Xbar
    S_out = (scan or solve_affine_scan_diag)(A, B, np.asarray(s0, dtype=np.float64))
    if damping != 1.0:
        S_out = (1 - damping) * S_in + damping * S_out
    return S_out, acc, la


def mala_rec(target, eps, Xi, U):
    """The taped MALA recursion's batched callbacks, as bound by
    src/interface.jl:409-468 (fwd_and_jvp_batch) for a target with batched
    logdensity/grad/hvp."""
    def fwd(Xbar, Z):
        fj = mala_step_batched_fwd_and_jvp(target, Xbar, eps, Xi, U, Z)
        return fj.FT, fj.Jt, fj.acc, fj.logalpha

    def jdiag(Xbar):
        return mala_jac_diag_batched(target, Xbar, eps, Xi, U)

    def jfull(Xbar, sigmoid=False):
        return mala_jac_full_batched(target, Xbar, eps, Xi, U, sigmoid)

    fwd.jac_full = jfull
    return fwd, jdiag


def deer_update(target, eps, Xi, U, s0, S_in, jacobian="stoch_diag", damping=1.0, probes=1,
                seed=0, it=0, scan=None):
    """deer_update_rec on the taped MALA recursion of ``target``."""
    fwd, jdiag = mala_rec(target, eps, Xi, U)
    return deer_update_rec(fwd, s0, S_in, jacobian, damping, probes, seed, it, scan, jdiag, fwd.jac_full)


def deer_solve_rec(fwd_and_jvp_batch, s0, T, tol_abs=1e-4, tol_rel=1e-3, maxiter=200,
                   jacobian="stoch_diag", damping=1.0, probes=1, seed=0, init=None, scan=None,
                   jac_diag_batch=None, jac_full_batch=None):
    """Newton iterations to the fixed point S = f(S).  src/DEER/DEER.jl:506-595.

    S starts as s0 in every column (:538-541).  metric = max|S_new - S| /
    (tol_abs + tol_rel max|S_new|); converged when metric <= 1 (:567-577);
    otherwise the last iterate is returned.  ``info.accepted`` is the hard
    accept mask of the final iteration (an output the reference does not
    expose)."""
    if maxiter < 1:
        raise ValueError("maxiter must be >= 1")
    if tol_abs < 0:
        raise ValueError("tol_abs must be >= 0")
    if tol_rel < 0:
        raise ValueError("tol_rel must be >= 0")
    if not (0 < damping <= 1):
        raise ValueError("damping must be in (0,1]")
    s0 = np.asarray(s0, dtype=np.float64)
    D = len(s0)
    if init is None:
        S = np.repeat(s0[:, None], T, axis=1)
    else:
        if init.shape != (D, T):
            raise ValueError("init must be size (D,T)")
        S = np.asarray(init, dtype=np.float64).copy()
    converged, metric, iters = False, math.inf, 0
    metrics = []
    acc = la = None
    for it in range(maxiter):
        iters = it + 1
        S_new, acc, la = deer_update_rec(fwd_and_jvp_batch, s0, S, jacobian, damping, probes, seed,
                                         it, scan, jac_diag_batch, jac_full_batch)
        dmax = float(np.max(np.abs(S_new - S))) if S.size else 0.0
        smax = float(np.max(np.abs(S_new))) if S.size else 0.0
        metric = dmax / (tol_abs + tol_rel * smax)
        metrics.append(metric)
        S = S_new
        if metric <= 1:
            converged = True
            break
    return S, DEERInfo(converged, iters, metric, jacobian, damping, probes, acc, la, metrics)


def deer_solve(target, eps, Xi, U, s0, tol_abs=1e-4, tol_rel=1e-3, maxiter=200,
               jacobian="stoch_diag", damping=1.0, probes=1, seed=0, init=None, scan=None):
    """deer_solve_rec on the taped MALA recursion of ``target``."""
    fwd, jdiag = mala_rec(target, eps, Xi, U)
    return deer_solve_rec(fwd, s0, Xi.shape[1], tol_abs, tol_rel, maxiter, jacobian, damping,
                          probes, seed, init, scan, jdiag, fwd.jac_full)


def trace_error(S, S_ref):
    """max|S - S_ref| / max|S_ref| (the metric form of src/DEER/DEER.jl:567-571)."""
    return float(np.max(np.abs(S - S_ref)) / max(np.max(np.abs(S_ref)), 1e-300))


# ----------------------------------------------------------------------------
# Adjoint of the taped recursion (no reference counterpart: the reference
# differentiates sequential MALA with Enzyme, ext/EnzymeExt.jl).
# ----------------------------------------------------------------------------


def run_mala_frozen(target, x0, eps, Xi, acc):
    """The taped MALA recursion with the accept decisions held at ``acc``:
    x_t = acc_t ? x_{t-1} + eps grad(x_{t-1}) + sqrt(2 eps) xi_t : x_{t-1}
    (the proposal of src/MALA/MALA.jl:207-229).  Returns D x T."""
    D, T = Xi.shape
    S = np.empty((D, T))
    x = np.asarray(x0, dtype=np.float64).copy()
    for t in range(T):
        if acc[t]:
            x = x + eps * target.grad(x) + math.sqrt(2 * eps) * Xi[:, t]
        S[:, t] = x
    return S


def mala_adjoint(target, eps, Xi, s0, S, acc, G):
    """Gradient of a trajectory statistic L(S) through the taped recursion at
    the trajectory S with the decisions ``acc`` held fixed.

    With Xbar_t = [s0, S][:, t] and J_t = a_t (I + eps H(Xbar_t)) + (1 - a_t) I:
      lambda_{T-1} = G_{T-1},  lambda_t = G_t + J_{t+1}^T lambda_{t+1},
      dL/deps   = sum_t a_t lambda_t . (grad(Xbar_t) + xi_t / sqrt(2 eps)),
      dL/dhyper = sum_t a_t eps lambda_t . dgrad/dhyper(Xbar_t),
      dL/ds0    = J_0^T lambda_0,
    where G = dL/dS.  H is symmetric, so J^T v = v + a eps H v.
    Returns (dL/deps, dL/dhyper (0 for targets without one), dL/ds0, lambda)."""
    D, T = S.shape
    Xbar = _xbar(s0, S)
    lam = np.zeros((D, T))
    carry = np.zeros(D)
    for t in range(T - 1, -1, -1):
        lam[:, t] = G[:, t] + carry
        carry = lam[:, t] + (eps * target.hvp(Xbar[:, t], lam[:, t]) if acc[t] else 0.0)
    a = np.asarray(acc, dtype=np.float64)[None, :]
    gx = target.grad_batch(Xbar)
    deps = float(np.sum(a * lam * (gx + Xi / math.sqrt(2 * eps))))
    dh = 0.0
    if hasattr(target, "dgrad_dhyper"):
        dh = float(np.sum(a * eps * lam * target.dgrad_dhyper(Xbar)))
    return deps, dh, carry, lam


# ----------------------------------------------------------------------------
# Tapes.
# ----------------------------------------------------------------------------


def make_tape(rng, D, T, dtype=np.float64):
    """xi_t ~ N(0, I), u_t ~ U(0,1) (src/interface.jl:334-358 draws them
    interleaved from Julia's RNG; here numpy's generator, in bulk)."""
    Xi = rng.standard_normal((T, D)).T.astype(dtype)
    U = rng.random(T).astype(dtype)
    return np.asfortranarray(Xi), U


def philox4x64_10(ctr, key):
    """Philox4x64-10 (Salmon et al., SC 2011) on Python integers: ten rounds
    of two 64x64->128 multiplies with the published constants and Weyl key
    schedule.  No reference counterpart (the reference draws its tape from
    Julia's RNG); pinned against numpy.random.Philox."""
    M0, M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
    W0, W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
    mask = (1 << 64) - 1
    x = [int(v) & mask for v in ctr]
    k0, k1 = int(key[0]) & mask, int(key[1]) & mask
    for _ in range(10):
        p0, p1 = M0 * x[0], M1 * x[2]
        x = [(p1 >> 64) ^ x[1] ^ k0, p1 & mask, (p0 >> 64) ^ x[3] ^ k1, p0 & mask]
        k0, k1 = (k0 + W0) & mask, (k1 + W1) & mask
    return x


def philox_words(seed, stream, ncalls):
    """Words of calls 0..ncalls-1 of Philox4x64-10 with key (seed, stream) and
    counter (j, 0, 0, 0), in call order: numpy's Philox started one step
    before counter 0 (it advances the counter before each call)."""
    m = np.iinfo(np.uint64).max
    bg = np.random.Philox(key=np.array([seed, stream], dtype=np.uint64),
                          counter=np.array([m, m, m, m], dtype=np.uint64))
    return bg.random_raw(4 * ncalls).astype(np.uint64)


def device_tape(seed, D, T, dtype=np.float64):
    """The opt-in on-device tape (no reference counterpart; it replaces the
    host draw of src/interface.jl:334-358 with another stream): U from
    stream 1 as ((w >> 11) + 1/2) 2^-53; Xi (column-major element i = tD + d)
    by Box-Muller on stream 0's word pairs (cosine for even i, sine for odd)."""
    n = D * T
    w = philox_words(seed, 1, (T + 3) // 4)[:T]
    U = ((w >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0**-53
    npair = (n + 1) // 2
    z = philox_words(seed, 0, (n + 3) // 4)[:2 * npair].reshape(npair, 2)
    u1 = ((z[:, 0] >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0**-53
    u2 = (z[:, 1] >> np.uint64(11)).astype(np.float64) * 2.0**-53
    r = np.sqrt(-2.0 * np.log(u1))
    th = 6.283185307179586 * u2
    flat = np.stack([r * np.cos(th), r * np.sin(th)], axis=1).reshape(-1)[:n]
    Xi = flat.reshape(T, D).T
    if dtype == np.float32:
        return np.asfortranarray(Xi.astype(np.float32)), np.minimum(U, 1.0 - 2.0**-24).astype(np.float32)
    return np.asfortranarray(Xi), U


# How each oracle function is pinned (tests/test_oracle_*.py).
PINNED = {
    "logq_mala_mass": "equals scipy.stats.multivariate_normal(x + eps M g, 2 eps M).logpdf(y) at 1e-12",
    "mala_step_taped_mass": "L = I reduces to mala_step_taped; forced accept/reject via u; the proposal's "
                            "mean and covariance (eps M g, 2 eps M) from 2e5 draws",
    "run_mala_sequential_taped_mass": "composition of mala_step_taped_mass; L = I equals run_mala_sequential_taped",
    "philox4x64_10": "equals numpy.random.Philox word for word (several keys and counters, including carries)",
    "device_tape": "Philox words from numpy's generator; U moments and Xi moments / normality (KS) on 2^18 draws",
    "run_mala_frozen": "equals run_mala_sequential_taped when fed that run's own decisions",
    "mala_sampler_run": "equals run_mala_sequential_taped on the tape drawn from the same RNG stream in the "
                        "same order (xi then u per step); logp of every state equals logp_batch of the states",
    "mala_adjoint": "central differences of L(run_mala_frozen) in eps, prior_scale and s0; "
                    "hard-decision differences where no decision flips",
    "logq_mala": "closed-form Gaussian log-density (scipy.stats.multivariate_normal) at 1e-12",
    "mala_step_taped": "forced accept/reject via u=tau/2 and (tau+1)/2; rejected => unchanged; determinism",
    "run_mala_sequential_taped": "stationary moments on N(0,I); DEER fixed point equals it; equals the "
                                 "independent C restatement (oracle/mala_seq.c) on every target: identical "
                                 "decisions, states to 1e-12; committed fixtures tests/golden/mala_seq_*.npz",
    "mala_step_batched_fwd_and_jvp": "column-by-column equals the scalar step+JVP; JVP equals central FD of the frozen-accept surrogate",
    "mala_jac_diag_batched": "equals jac_diag_via_jvps (D basis JVPs) at 1e-10",
    "targets": "grad = central FD of logp, hvp = central FD of grad (FP64), logreg vs explicit sums",
    "solve_affine_scan_diag": "equals the sequential recurrence at 1e-12 on the reference's shape list; residual known answers 0.25 / 0.5; T=0",
    "deer_update": "batched Hutchinson is exact on a diagonal-Jacobian recursion (z*z=1)",
    "deer_solve": "fixed point equals run_mala_sequential_taped (decisions identical); defect <= 1e-9; converges in <= T undamped iterations",
    "mala_jac_full_batched": "equals jac_full_via_jvps (D basis JVPs) at 1e-10; full-Jacobian DEER reaches the sequential trajectory in few iterations",
    "rademacher": "entries are +-1; mean ~0 and pairwise decorrelation over 1e6 draws",
}
