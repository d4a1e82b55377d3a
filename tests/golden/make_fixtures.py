"""Writes the sequential-MALA fixtures of this directory.

Each ``mala_seq_<target>.npz`` holds a small seeded problem (target data, eps,
x0, tape Xi/U) and the sequential taped MALA result (states S, decisions acc,
log alpha la) computed in FP64 by the C restatement (oracle/mala_seq.c),
checked here against the numpy restatement (oracle/deer_oracle.py): identical
decisions, states within 1e-12.  The reference itself (Julia) is not run.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import deer_oracle as O  # noqa: E402
from oracle import mala_seq  # noqa: E402


def problems():
    rng = np.random.default_rng(20261204)
    out = {}
    D, T = 8, 256
    prec = rng.uniform(0.5, 3.0, D)
    mean = rng.standard_normal(D)
    out["diag"] = (dict(prec=prec, mean=mean), O.DiagGaussian(prec, mean), 0.25, D, T)
    D, T = 6, 256
    Lam, ev = O.make_corr_gaussian(rng, D, 2.0)
    mu = rng.standard_normal(D)
    out["gauss"] = (dict(Lam=Lam, mean=mu), O.Gaussian(Lam, mu), 0.9 / ev.max(), D, T)
    D, T, N = 5, 256, 64
    X, y, _ = O.make_logreg_data(rng, N, D)
    out["logreg"] = (dict(X=X, y=y, prior_scale=1.5), O.LogReg(X, y, 1.5), 0.15, D, T)
    for kind, (kw, tgt, eps, D, T) in out.items():
        x0 = rng.standard_normal(D)
        Xi = rng.standard_normal((D, T))
        U = rng.random(T)
        out[kind] = (kw, tgt, eps, x0, Xi, U)
    return out


def main():
    for kind, (kw, tgt, eps, x0, Xi, U) in problems().items():
        S, acc, la = mala_seq.run(kind, x0, eps, Xi, U, **kw)
        S2, acc2, la2 = O.run_mala_sequential_taped(tgt, x0, eps, Xi, U)
        assert np.array_equal(acc, acc2), kind
        assert np.max(np.abs(S - S2)) <= 1e-12 * max(1.0, np.max(np.abs(S))), kind
        data = {k: np.asarray(v, dtype=np.float64) for k, v in kw.items()}
        np.savez_compressed(os.path.join(HERE, f"mala_seq_{kind}.npz"), eps=np.float64(eps), x0=x0, Xi=Xi, U=U,
                            S=S, acc=acc, la=la, **data)
        print(kind, "accepted", int(acc.sum()), "of", len(acc))


if __name__ == "__main__":
    main()
