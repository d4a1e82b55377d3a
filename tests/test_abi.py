"""CPU checks of the C ABI boundary: the library loads and exports every
entry point declared in include/pmcmc_amd.h; host-side argument validation
mirrors the reference's ArgumentError / DimensionMismatch behaviour."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pmcmc_amd.h")
LIB = os.path.join(ROOT, "parallelmcmc.jl_amd", "libpmcmc_amd.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmcmc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first: python parallelmcmc.jl_amd/build_lib.py"
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from ParallelMCMC import _lib

    assert set(_lib.exported_symbols()) == set(declared_symbols())


def test_version_string_without_gpu():
    from ParallelMCMC import _lib

    assert b"gfx950" in _lib.load().pmcmc_version()


def test_null_context_is_rejected_without_gpu():
    from ParallelMCMC import _lib

    lib = _lib.load()
    assert lib.pmcmc_ctx_create(0, None, None) == _lib.PMCMC_ERR_ARGUMENT
    assert lib.pmcmc_scan_diag_f32(None, None, None, None, None, 1, 1) == _lib.PMCMC_ERR_ARGUMENT
    assert lib.pmcmc_ctx_destroy(None) == _lib.PMCMC_OK


def test_sampler_argument_validation():
    # src/interface.jl:273-275 and DEER.jl:525-528 checks, host side
    from ParallelMCMC import ArgumentError, DEER, DensityModel, ParallelMALASampler, StdNormalTarget

    with pytest.raises(ArgumentError):
        ParallelMALASampler(0.0)
    with pytest.raises(ArgumentError):
        ParallelMALASampler(0.1, jacobian="bogus")
    assert ParallelMALASampler(0.1, jacobian=":full").jacobian == "full"
    s = ParallelMALASampler(0.1)
    assert (s.T, s.maxiter, s.tol_abs, s.tol_rel, s.jacobian, s.damping, s.probes) == (64, 200, 1e-6, 1e-5, "stoch_diag", 0.5, 1)
    model = DensityModel(dim=3, target=StdNormalTarget(3))
    rec = DEER.TapedRecursion(model, 0.1, np.zeros((3, 4)), np.full(4, 0.5))
    for kw in (dict(maxiter=0), dict(tol_abs=-1.0), dict(tol_rel=-1.0), dict(damping=0.0), dict(probes=0),
               dict(jacobian="bogus"), dict(jacobian="full")):  # :full needs a dense Gaussian target
        with pytest.raises(ArgumentError):
            DEER.solve(rec, np.zeros(3), **kw)
    with pytest.raises(ArgumentError):
        DEER.solve(rec, np.zeros(2))
    with pytest.raises(ArgumentError):  # closures only: no GPU target descriptor
        DEER.TapedRecursion(DensityModel(lambda x: 0.0, lambda x: x, 3), 0.1, np.zeros((3, 4)), np.full(4, 0.5))


def test_full_jacobian_host_validation():
    """jacobian=:full needs a dense Gaussian target with D <= 64 (host-side checks)."""
    from ParallelMCMC import ArgumentError, DEER, DensityModel, GaussianTarget, StdNormalTarget

    rec = DEER.TapedRecursion(DensityModel(dim=3, target=StdNormalTarget(3)), 0.1, np.zeros((3, 4)),
                              np.full(4, 0.5))
    with pytest.raises(ArgumentError):
        DEER.solve(rec, np.zeros(3), jacobian="full")
    rec = DEER.TapedRecursion(DensityModel(dim=65, target=GaussianTarget(np.eye(65))), 0.1, np.zeros((65, 4)),
                              np.full(4, 0.5))
    with pytest.raises(ArgumentError):
        DEER.solve(rec, np.zeros(65), jacobian="full")
