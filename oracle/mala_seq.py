"""ctypes loader of ``mala_seq.c``, the plain-C restatement of sequential taped
MALA (FP64) used to cross-check the numpy oracle.

TEST INFRASTRUCTURE ONLY: loaded by ``tests/`` and ``tests/golden/``'s fixture
script as a checker, never by the product path.  ``build()`` compiles it with
gcc into ``oracle/libmala_seq.so``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "mala_seq.c")
LIB = os.path.join(HERE, "libmala_seq.so")

KIND = {"diag": 0, "gauss": 1, "logreg": 2}


class _Target(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("D", C.c_int64), ("N", C.c_int64),
                ("prec", C.c_void_p), ("mean", C.c_void_p), ("Lam", C.c_void_p), ("X", C.c_void_p),
                ("y", C.c_void_p), ("prior_scale", C.c_double)]


def build() -> str:
    """Compile the C restatement (always recompiles)."""
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O2", "-std=c11", "-D_DEFAULT_SOURCE", "-fPIC", "-shared", "-o", tmp, SRC, "-lm"],
                   check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.mala_seq_taped.restype = C.c_int
        lib.mala_seq_taped.argtypes = [C.POINTER(_Target), C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = lib
    return _lib


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def run(kind, x0, eps, Xi, U, prec=None, mean=None, Lam=None, X=None, y=None, prior_scale=1.0):
    """Sequential taped MALA with the C restatement: returns (S, accepted, logalpha)
    like ``deer_oracle.run_mala_sequential_taped``.  ``kind``: 'diag', 'gauss'
    or 'logreg'; X is the N x D design."""
    lib = load()
    keep = []

    def ptr(a, order="F"):
        if a is None:
            return None
        b = np.asarray(a, dtype=np.float64, order=order)
        b = np.require(b, requirements=["F_CONTIGUOUS"] if order == "F" else ["C_CONTIGUOUS"])
        keep.append(b)
        return b.ctypes.data

    Xi = np.asarray(Xi, dtype=np.float64, order="F")
    D, T = Xi.shape
    t = _Target()
    t.kind = KIND[kind]
    t.D = D
    t.N = 0 if X is None else np.asarray(X).shape[0]
    t.prec = ptr(prec)
    t.mean = ptr(mean)
    t.Lam = ptr(Lam)
    t.X = ptr(X)
    t.y = ptr(y)
    t.prior_scale = float(prior_scale)
    S = np.empty((D, T), dtype=np.float64, order="F")
    acc = np.empty(T, dtype=np.uint8)
    la = np.empty(T, dtype=np.float64)
    x0_ = _f64(x0)
    U_ = _f64(U)
    Xi_ = np.require(Xi, requirements=["F_CONTIGUOUS"])
    rc = lib.mala_seq_taped(C.byref(t), float(eps), x0_.ctypes.data, Xi_.ctypes.data, U_.ctypes.data, T,
                            S.ctypes.data, acc.ctypes.data, la.ctypes.data)
    if rc != 0:
        raise ValueError("mala_seq_taped: bad arguments")
    return S, acc.astype(bool), la
