"""GPU tests of the opt-in on-device tape (pmcmc_tape_generate; no reference
counterpart: the reference draws the tape on the host,
src/interface.jl:334-358) against the oracle's restatement (device_tape),
whose generator words come from numpy.random.Philox:

* FP64: U bit for bit (the Philox words decide it exactly), Xi within 1e-14
  relative (Box-Muller through the device's log / sincos);
* FP32: the rounded FP64 tape (U exact, Xi within one FP32 ulp);
* ragged sizes (D T not a multiple of 4) and a large block;
* sample() with tape_rng="device": every block is sequential MALA on the
  oracle's tape for the seed drawn from the same Generator.
"""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,D,T", [(1, 1, 1), (7, 3, 5), (2**63 + 11, 13, 101), (42, 100, 4096)])
def test_device_tape_f64_matches_oracle(seed, D, T):
    from ParallelMCMC import _lib
    from ParallelMCMC.interface import make_mala_tape_block_device

    Xi, U = make_mala_tape_block_device(seed, D, T, _lib.PREC_F64)
    Xr, Ur = O.device_tape(seed, D, T)
    Xi, U = Xi.cpu().numpy(), U.cpu().numpy()
    assert np.array_equal(U, Ur)
    assert np.max(np.abs(Xi - Xr)) <= 1e-14 * max(1.0, np.max(np.abs(Xr)))


def test_device_tape_f32_matches_oracle():
    from ParallelMCMC import _lib
    from ParallelMCMC.interface import make_mala_tape_block_device

    D, T = 37, 999
    Xi, U = make_mala_tape_block_device(3, D, T, _lib.PREC_F32)
    Xr, Ur = O.device_tape(3, D, T, np.float32)
    Xi, U = Xi.cpu().numpy(), U.cpu().numpy()
    assert Xi.dtype == np.float32
    assert np.array_equal(U, Ur)
    ulp = np.spacing(np.abs(Xr))
    assert np.all(np.abs(Xi - Xr) <= ulp)


def test_sampler_with_device_tape_matches_sequential():
    from ParallelMCMC import DensityModel, LogisticTarget, ParallelMALASampler, sample

    rng0 = np.random.default_rng(9)
    Xd, y, _ = O.make_logreg_data(rng0, 200, 5)
    tgt = O.LogReg(Xd, y, 1.0)
    model = DensityModel(dim=5, target=LogisticTarget(Xd, y, 1.0))
    D, N, T, eps = 5, 100, 32, 0.03
    sampler = ParallelMALASampler(eps, T=T, tol_abs=1e-13, tol_rel=1e-12, tape_rng="device")
    samples = sample(np.random.default_rng(5), model, sampler, N)
    got = samples.array()
    rng = np.random.default_rng(5)
    x = rng.standard_normal(D)
    row = 0
    while row < N:
        Xi, U = O.device_tape(int(rng.integers(0, 2**63 - 1)), D, T)
        rng.integers(0, 2**63 - 1)  # the block's probe seed
        ref, _, _ = O.run_mala_sequential_taped(tgt, x, eps, Xi, U)
        nkeep = min(T, N - row)
        assert np.max(np.abs(got[row:row + nkeep].T - ref[:, :nkeep])) < 1e-10
        x = got[row + nkeep - 1].copy()
        row += nkeep
