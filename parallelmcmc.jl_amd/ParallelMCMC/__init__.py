"""ParallelMCMC on MI355X: DEER's Newton solve over taped MALA trajectories,
run by hand-written gfx950 HIP kernels behind the C ABI of
include/pmcmc_amd.h.

Surface mirrored from the reference package (src/ParallelMCMC.jl:32-41):
``DensityModel``, ``ParallelMALASampler``, ``sample``, ``step``, the
``DEER``, ``DEERScan``, ``MALA`` submodules and the owned ``pmcmc_matmul``,
``pmcmc_dot``, ``pmcmc_dotsum``.
"""
from . import _lib
from ._lib import ArgumentError, DimensionMismatch, PmcmcError
from .interface import (
    DensityModel,
    ParallelMALABlockSamples,
    ParallelMALASampler,
    ParallelMALAState,
    ParallelMALATransition,
    sample,
    step,
)
from .targets import DiagGaussianTarget, GaussianTarget, LogisticTarget, StdNormalTarget
from . import DEER, DEERScan, MALA  # noqa: E402
from . import Adjoint, Batched, Tune  # noqa: E402
from .owned import pmcmc_dot, pmcmc_dotsum, pmcmc_matmul  # noqa: E402

__all__ = [
    "DensityModel",
    "ParallelMALASampler",
    "ParallelMALABlockSamples",
    "ParallelMALAState",
    "ParallelMALATransition",
    "sample",
    "step",
    "DEER",
    "DEERScan",
    "MALA",
    "Adjoint",
    "Batched",
    "Tune",
    "pmcmc_matmul",
    "pmcmc_dot",
    "pmcmc_dotsum",
    "DiagGaussianTarget",
    "StdNormalTarget",
    "LogisticTarget",
    "GaussianTarget",
    "ArgumentError",
    "DimensionMismatch",
    "PmcmcError",
]
