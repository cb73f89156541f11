"""Pins of the oracle's restatement of the int8 Fock contraction (oracle.core.int8_fock), which
the GPU kernel must match bit for bit (tests/test_gpu_int8.py):

* the digit split reproduces every 35-bit integer with digits in [-64, 64];
* operands that fit the first three digits give the exact product (every digit product then
  has level <= 6 and is kept), for either sign of the weights: a wrong level weight, a dropped
  level (the (3, 3) products are level 6) or a wrong sign fails it;
* random operands agree with FP64 to the 35-bit representation (a 28-bit, four-digit version
  misses this bound by two orders of magnitude);
* the chunking does not change the exact case."""
import numpy as np

from oracle import core


def test_split5_reconstructs_and_bounds_digits():
    rng = np.random.default_rng(1)
    J = np.concatenate([rng.integers(-2 ** 34, 2 ** 34 + 1, size=100000),
                        [2 ** 34, -2 ** 34, 0, 8191, -8192, 8192, 2 ** 20 * 16384]]).astype(np.int64)
    d = core._split5(J)
    rec = sum(x << (7 * (4 - k)) for k, x in enumerate(d))
    assert np.array_equal(rec, J)
    for x in d:
        assert x.min() >= -64 and x.max() <= 64


def _three_digit_system(rng, N, n, sign):
    # values k 2^-16 (|k| <= 2^16): at the chunk scale 2^0 they are J = k 2^18, digits d1..d3;
    # |wbw| = 4 (sqrt exact), v integers 1..8 with 8 present per chunk of 128: scB = 2^-4 and
    # JB = J v / 16 keeps the low 14 bits zero
    phi = rng.integers(-2 ** 16, 2 ** 16, size=(N, n)) * 2.0 ** -16
    v = rng.integers(1, 9, size=N).astype(np.float64)
    for c in range(0, N, 128):
        phi[c] = 0.5           # sqrt|wbw| phi = 1: the chunk maximum of every function
        v[c] = 8.0
    wbw = np.full(N, 4.0) * sign
    return phi, wbw, v


def test_int8_fock_exact_when_operands_fit_three_digits():
    rng = np.random.default_rng(2)
    for sign in (1.0, -1.0):
        phi, wbw, v = _three_digit_system(rng, 512, 7, sign)
        F = core.int8_fock(phi, wbw, v, 128)
        Fx = phi.T @ ((wbw * v)[:, None] * phi)
        assert np.array_equal(F, Fx), np.abs(F - Fx).max()
        F2 = core.int8_fock(phi, wbw, v, 256)    # other chunking, same exact sums
        assert np.array_equal(F2, Fx)


def test_int8_fock_random_operands_within_35_bits():
    rng = np.random.default_rng(3)
    N, n = 1024, 9
    phi = rng.normal(size=(N, n)) * np.exp(rng.normal(size=(N, 1)))
    wbw = rng.uniform(0.1, 2.0, size=N) * np.where(rng.random(N) < 0.1, -1.0, 1.0)
    v = rng.uniform(0.5, 30.0, size=N)
    F = core.int8_fock(phi, wbw, v, 256)
    Fx = phi.T @ ((wbw * v)[:, None] * phi)
    scale = np.abs(phi).T @ ((np.abs(wbw * v))[:, None] * np.abs(phi))
    rel = np.abs(F - Fx) / scale
    print('int8 oracle max error / sum|terms| %.3e' % rel.max())
    assert rel.max() < 2.0 ** -30
    assert np.array_equal(F, F.T)
