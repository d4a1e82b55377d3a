"""The two restatements of sequential taped MALA agree: numpy
(oracle/deer_oracle.py) and plain C (oracle/mala_seq.c), written
independently from src/MALA/MALA.jl:172-346 and the analytic targets.  On
seeded tapes for every target they must take identical accept decisions and
states equal to round-off; a plausibly wrong restatement (log q dropped from
log alpha, or the proposal's gradient step at the wrong point) disagrees."""
import math

import numpy as np
import pytest

from oracle import deer_oracle as O
from oracle import mala_seq


def _cases():
    rng = np.random.default_rng(11)
    out = []
    D = 6
    prec = rng.uniform(0.5, 3.0, D)
    mean = rng.standard_normal(D)
    out.append(("diag", O.DiagGaussian(prec, mean), dict(prec=prec, mean=mean), 0.3, D))
    Lam, ev = O.make_corr_gaussian(rng, 5, 2.0)
    mu = rng.standard_normal(5)
    out.append(("gauss", O.Gaussian(Lam, mu), dict(Lam=Lam, mean=mu), 0.8 / ev.max(), 5))
    X, y, _ = O.make_logreg_data(rng, 40, 4)
    out.append(("logreg", O.LogReg(X, y, 1.3), dict(X=X, y=y, prior_scale=1.3), 0.05, 4))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_c_and_numpy_restatements_agree(case):
    kind, tgt, kw, eps, D = case
    rng = np.random.default_rng({"diag": 1, "gauss": 2, "logreg": 3}[kind])
    T = 400
    x0 = rng.standard_normal(D)
    Xi = rng.standard_normal((D, T))
    U = rng.random(T)
    S_np, acc_np, la_np = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
    S_c, acc_c, la_c = mala_seq.run(kind, x0, eps, Xi, U, **kw)
    assert 0 < acc_np.sum() < T  # both branches exercised
    assert np.array_equal(acc_np, acc_c)
    assert np.max(np.abs(S_np - S_c)) <= 1e-12 * max(1.0, np.max(np.abs(S_np)))
    assert np.max(np.abs(la_np - la_c)) <= 1e-9


def test_zero_uniform_accepts_in_both():
    tgt = O.DiagGaussian.std_normal(3)
    x0 = np.array([5.0, -5.0, 5.0])  # far out: log alpha of a step towards the mode is large
    Xi = np.zeros((3, 2))
    U = np.array([0.0, 0.5])
    S_np, acc_np, _ = O.run_mala_sequential_taped(tgt, x0, 0.1, Xi, U, check_u=False)
    S_c, acc_c, _ = mala_seq.run("diag", x0, 0.1, Xi, U, prec=np.ones(3), mean=np.zeros(3))
    assert acc_np[0] and acc_c[0]
    assert np.array_equal(acc_np, acc_c) and np.allclose(S_np, S_c, atol=1e-14)


def _buggy_logalpha_run(tgt, x0, eps, Xi, U):
    """Planted mistake: log alpha without the proposal densities."""
    D, T = Xi.shape
    x = x0.copy()
    acc = np.zeros(T, dtype=bool)
    for t in range(T):
        y = x + eps * tgt.grad(x) + math.sqrt(2 * eps) * Xi[:, t]
        la = tgt.logp(y) - tgt.logp(x)
        acc[t] = math.log(U[t]) < la
        if acc[t]:
            x = y
    return acc


def test_planted_mistake_is_caught():
    kind, tgt, kw, eps, D = _cases()[2]  # logistic
    rng = np.random.default_rng(5)
    T = 300
    x0 = rng.standard_normal(D)
    Xi = rng.standard_normal((D, T))
    U = rng.random(T)
    _, acc_c, _ = mala_seq.run(kind, x0, eps, Xi, U, **kw)
    assert not np.array_equal(_buggy_logalpha_run(tgt, x0, eps, Xi, U), acc_c)
