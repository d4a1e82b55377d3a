"""Autotuner of the scan tile shapes (the library's ``pmcmc_scan_tune`` /
``pmcmc_deer_tune`` and its shape table).  No reference counterpart: the
reference's scan has no tile shapes (src/DEER/DEERScan.jl:102-173).

The level-0 tile of the diagonal scan and of the fused DEER tile kernels
(``tc`` columns per tile, ``parts`` parts per row of the in-tile scan, and
for the standalone scan the single-pass forms: 1 LDS-tiled look-back, 2
register-resident tiles with look-back for 16 <= D <= 256, 3 the same with
persistent workgroups) is looked up in a
process-wide table keyed by (kind, precision, D, ceil(log2 T)); without an
entry a fixed rule applies.  ``tune_scan`` measures candidate shapes of the
standalone diagonal scan on synthetic operands, ``tune_solve`` the
shape-dependent kernels of a few Newton iterations of a given problem (and
re-plans the workspace it used).  Both store the winner in the table, where
later scans, workspaces and batched solves pick it up.  ``save`` / ``load``
keep the table in a JSON file.
"""
from __future__ import annotations

import ctypes as C
import json

import torch

from . import _device, _lib
from ._lib import ArgumentError

KINDS = {"scan_diag": _lib.TUNE_SCAN_DIAG, "deer": _lib.TUNE_DEER}
_KIND_NAMES = {v: k for k, v in KINDS.items()}


def _prec_code(dtype):
    if dtype in (torch.float64, "f64", "float64"):
        return _lib.PREC_F64
    if dtype in (torch.float32, "f32", "float32", "bf16_jac"):
        return _lib.PREC_F32
    raise ArgumentError(f"unsupported precision {dtype!r}")


def _shape_dict(s: _lib.TileShape):
    return {"tc": int(s.tc), "parts": int(s.parts), "single_pass": int(s.single_pass), "us": float(s.us)}


def tune_scan(D: int, T: int, dtype=torch.float32, reps: int = 5, device=None) -> dict:
    """Time candidate tile shapes of the diagonal scan of a D x T problem and
    keep the fastest (returned with its time per scan in microseconds)."""
    dev = device if device is not None else torch.cuda.current_device()
    ctx = _device.context(dev)
    best = _lib.TileShape()
    ctx.call("pmcmc_scan_tune", _prec_code(dtype), int(D), int(T), int(reps), C.byref(best))
    return _shape_dict(best)


def tune_solve(rec, s0, iters: int = 6, workspace=None, jacobian="stoch_diag", probes: int = 1, seed: int = 0,
               damping: float = 1.0) -> dict:
    """Time candidate tile shapes on ``iters`` Newton iterations of the taped
    recursion ``rec`` from ``s0`` (tile reduce, spine and tile apply only, the
    kernels the shape changes) and keep the fastest; ``workspace``, when given,
    is re-planned with it (new workspaces of this size pick it up anyway)."""
    from . import DEER

    jcode, jname = DEER._jac_code(jacobian)
    D, T = rec.D, rec.T
    ws = workspace if workspace is not None else DEER.DEERWorkspace(D, T, rec.precision)
    if (ws.D, ws.T) != (D, T) or ws.precision != rec.precision:
        raise ArgumentError("workspace does not match the recursion")
    dt = _device.torch_dtype(rec.precision)
    Xi, U = rec.device_tape(ws.device)
    x0 = _device.to_vec(s0, dt, ws.device)
    params = _lib.DeerParams(float(rec.eps), 0.0, 0.0, float(damping), int(iters), jcode, int(probes),
                             rec.precision, int(seed) & 0xFFFFFFFFFFFFFFFF, 0)
    best = _lib.TileShape()
    ws.ctx.call("pmcmc_deer_tune", ws.h, rec.target.handle(ws.ctx), C.byref(params), _device.ptr(x0),
                _device.ptr(Xi), _device.ptr(U), int(iters), C.byref(best))
    return _shape_dict(best)


def lookup(kind: str, dtype, D: int, T: int):
    """The table's shape for (kind, precision, D, T), or None."""
    out = _lib.TileShape()
    _lib.check(_lib.load().pmcmc_tune_get(KINDS[kind], _prec_code(dtype), int(D), int(T), C.byref(out)), None,
               "pmcmc_tune_get")
    return _shape_dict(out) if out.tc > 0 else None


def set_entry(kind: str, dtype, D: int, T: int, tc: int, parts: int, single_pass: int = 0, us: float = 0.0):
    """Put a shape in the table (raises if the tile kernels cannot run it).
    ``single_pass``: 0 two-pass, 1 LDS-tiled look-back, 2 register tiles,
    3 persistent register tiles (both for 16 <= D <= 256, run as row blocks of
    at most 64 rows; ``tc`` = 64 or 128 columns, or 256 in FP32 one tile per
    workgroup)."""
    s = _lib.TileShape(int(tc), int(parts), int(single_pass), 0, float(us))
    rc = _lib.load().pmcmc_tune_set(KINDS[kind], _prec_code(dtype), int(D), int(T), C.byref(s))
    if rc != 0:
        raise ArgumentError(f"tile shape tc={tc} parts={parts} does not fit D={D} ({kind})")


def entries() -> list:
    """Every table entry as a dict (T is given as its bin, ceil(log2 T))."""
    lib = _lib.load()
    out = []
    for i in range(int(lib.pmcmc_tune_count())):
        kind, prec, log2T = C.c_int(), C.c_int(), C.c_int32()
        D = C.c_int64()
        s = _lib.TileShape()
        _lib.check(lib.pmcmc_tune_entry(i, C.byref(kind), C.byref(prec), C.byref(D), C.byref(log2T), C.byref(s)),
                   None, "pmcmc_tune_entry")
        e = _shape_dict(s)
        e.update(kind=_KIND_NAMES[kind.value], precision="f64" if prec.value == _lib.PREC_F64 else "f32",
                 D=int(D.value), log2T=int(log2T.value))
        out.append(e)
    return out


def clear():
    _lib.load().pmcmc_tune_clear()


def save(path: str):
    with open(path, "w") as f:
        json.dump(entries(), f, indent=1)


def load(path: str):
    """Merge a saved table into the current one."""
    with open(path) as f:
        for e in json.load(f):
            set_entry(e["kind"], e["precision"], e["D"], 1 << int(e["log2T"]), e["tc"], e["parts"],
                      int(e.get("single_pass", 0)), e.get("us", 0.0))
