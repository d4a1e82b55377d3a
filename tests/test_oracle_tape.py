"""Pins of the oracle's opt-in device-tape restatement (no reference
counterpart: the reference draws its tape from Julia's RNG,
src/interface.jl:334-358).

* philox4x64_10 (Python integers) equals numpy.random.Philox word for word,
  for several keys and counters, including carries between counter words;
* philox_words (numpy's generator started one step before counter 0) gives
  call j at counter (j, 0, 0, 0);
* device_tape: U in (0, 1) with the uniform's moments; Xi standard normal
  (moments, Kolmogorov-Smirnov) on 2^18 draws; the FP32 tape is the rounded
  FP64 tape with U below 1.
"""
import numpy as np
import pytest
from scipy import stats

from oracle import deer_oracle as O


@pytest.mark.parametrize("key,ctr", [((0, 0), (0, 0, 0, 0)), ((123, 456), (5, 0, 0, 0)),
                                     ((2**64 - 1, 17), (2**64 - 1, 3, 0, 0)),
                                     ((0x243F6A8885A308D3, 0x13198A2E03707344), (7, 2**64 - 1, 2**64 - 1, 9))])
def test_philox_matches_numpy(key, ctr):
    m = (1 << 64) - 1
    # numpy advances the 256-bit counter before each call: start one below
    c = (ctr[0] | ctr[1] << 64 | ctr[2] << 128 | ctr[3] << 192) - 1
    c &= (1 << 256) - 1
    start = np.array([(c >> (64 * i)) & m for i in range(4)], dtype=np.uint64)
    bg = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=start)
    got = [int(v) for v in bg.random_raw(8)]
    assert got[:4] == O.philox4x64_10(ctr, key)
    c1 = (c + 2) & ((1 << 256) - 1)
    nxt = [(c1 >> (64 * i)) & m for i in range(4)]
    assert got[4:] == O.philox4x64_10(nxt, key)


def test_philox_words_counter_layout():
    w = O.philox_words(99, 1, 3)
    for j in range(3):
        assert [int(v) for v in w[4 * j:4 * j + 4]] == O.philox4x64_10((j, 0, 0, 0), (99, 1))


def test_device_tape_distribution():
    D, T = 64, 4096
    Xi, U = O.device_tape(2024, D, T)
    assert Xi.shape == (D, T) and U.shape == (T,)
    assert np.all(U > 0) and np.all(U < 1)
    assert abs(U.mean() - 0.5) < 0.02 and abs(U.var() - 1 / 12) < 0.01
    z = Xi.reshape(-1)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # the streams differ with the seed, and the noise's cos/sin pairs are uncorrelated
    Xi2, U2 = O.device_tape(2025, D, T)
    assert not np.array_equal(U, U2)
    assert abs(np.corrcoef(z[0::2], z[1::2])[0, 1]) < 0.02


def test_device_tape_f32_is_rounded_f64():
    Xi, U = O.device_tape(5, 7, 33)
    Xf, Uf = O.device_tape(5, 7, 33, np.float32)
    assert Xf.dtype == np.float32 and Uf.dtype == np.float32
    assert np.array_equal(Xf, Xi.astype(np.float32))
    assert np.all(Uf < 1) and np.all(Uf > 0)
