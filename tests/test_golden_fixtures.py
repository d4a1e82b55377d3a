"""The committed sequential-MALA fixtures (tests/golden/mala_seq_*.npz, made by
tests/golden/make_fixtures.py) are reproduced by both restatements of the
oracle: identical decisions, states and log alpha to round-off."""
import os

import numpy as np
import pytest

from oracle import deer_oracle as O
from oracle import mala_seq

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KINDS = ["diag", "gauss", "logreg"]


def load(kind):
    with np.load(os.path.join(GOLDEN, f"mala_seq_{kind}.npz")) as f:
        return {k: f[k] for k in f.files}


def numpy_target(kind, f):
    if kind == "diag":
        return O.DiagGaussian(f["prec"], f["mean"])
    if kind == "gauss":
        return O.Gaussian(f["Lam"], f["mean"])
    return O.LogReg(f["X"], f["y"], float(f["prior_scale"]))


def c_kwargs(kind, f):
    if kind == "diag":
        return dict(prec=f["prec"], mean=f["mean"])
    if kind == "gauss":
        return dict(Lam=f["Lam"], mean=f["mean"])
    return dict(X=f["X"], y=f["y"], prior_scale=float(f["prior_scale"]))


@pytest.mark.parametrize("kind", KINDS)
def test_both_oracles_reproduce_fixture(kind):
    f = load(kind)
    assert 0 < f["acc"].sum() < len(f["acc"])
    for S, acc, la in (O.run_mala_sequential_taped(numpy_target(kind, f), f["x0"], float(f["eps"]), f["Xi"], f["U"]),
                       mala_seq.run(kind, f["x0"], float(f["eps"]), f["Xi"], f["U"], **c_kwargs(kind, f))):
        assert np.array_equal(acc, f["acc"])
        assert np.max(np.abs(S - f["S"])) <= 1e-12 * max(1.0, np.max(np.abs(f["S"])))
        assert np.max(np.abs(la - f["la"])) <= 1e-9
