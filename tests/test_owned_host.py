"""Host-side checks of the owned matmul / dot / dotsum wrappers (no GPU):
shape errors raise DimensionMismatch and unsupported element types or host
tensors raise ArgumentError before anything reaches the library, as the
reference's `A * B`, `dot`, `sum(A .* B)` would (src/ParallelMCMC.jl:23-25)."""
import pytest
import torch

from ParallelMCMC import ArgumentError, DimensionMismatch, pmcmc_dot, pmcmc_dotsum, pmcmc_matmul


def test_matmul_inner_dimension_mismatch():
    with pytest.raises(DimensionMismatch):
        pmcmc_matmul(torch.zeros(4, 3, dtype=torch.float64), torch.zeros(2, 5, dtype=torch.float64))
    with pytest.raises(DimensionMismatch):
        pmcmc_matmul(torch.zeros(4, 3, dtype=torch.float64), torch.zeros(4, dtype=torch.float64))


def test_matmul_rejects_vector_times_vector_and_other_ranks():
    with pytest.raises(ArgumentError):
        pmcmc_matmul(torch.zeros(3, dtype=torch.float64), torch.zeros(3, dtype=torch.float64))
    with pytest.raises(ArgumentError):
        pmcmc_matmul(torch.zeros(2, 3, 4, dtype=torch.float64), torch.zeros(4, 2, dtype=torch.float64))


def test_element_types_and_devices():
    with pytest.raises(ArgumentError):
        pmcmc_matmul(torch.zeros(4, 3, dtype=torch.int64), torch.zeros(3, 5, dtype=torch.int64))
    with pytest.raises(ArgumentError):
        pmcmc_matmul(torch.zeros(4, 3, dtype=torch.float32), torch.zeros(3, 5, dtype=torch.float64))
    with pytest.raises(ArgumentError):  # host tensors: the kernels need device memory
        pmcmc_matmul(torch.zeros(4, 3, dtype=torch.float64), torch.zeros(3, 5, dtype=torch.float64))


def test_dot_and_dotsum_shapes():
    with pytest.raises(DimensionMismatch):
        pmcmc_dot(torch.zeros(3, dtype=torch.float64), torch.zeros(4, dtype=torch.float64))
    with pytest.raises(ArgumentError):
        pmcmc_dot(torch.zeros(3, 1, dtype=torch.float64), torch.zeros(3, 1, dtype=torch.float64))
    with pytest.raises(DimensionMismatch):
        pmcmc_dotsum(torch.zeros(3, 2, dtype=torch.float64), torch.zeros(2, 3, dtype=torch.float64))


def test_sampler_rejects_bad_mass_factor():
    """cholM must be a square lower-triangular factor with a positive diagonal."""
    import numpy as np
    import pytest

    from ParallelMCMC import ArgumentError, ParallelMALASampler

    with pytest.raises(ArgumentError):
        ParallelMALASampler(0.1, cholM=np.ones((3, 3)))
    with pytest.raises(ArgumentError):
        ParallelMALASampler(0.1, cholM=-np.eye(3))
    with pytest.raises(ArgumentError):
        ParallelMALASampler(0.1, cholM=np.ones((3, 2)))
    ParallelMALASampler(0.1, cholM=np.eye(3))
