"""Pins of the oracle's restatement of the int8 Fock contraction (oracle.core.int8_fock), which
the GPU kernel must match bit for bit (tests/test_gpu_int8.py):

* the digit split reproduces every 35-bit integer with digits in [-64, 64];
* operands that fit the first three digits give the exact product (every digit product then
  has level <= 6 and is kept), for either sign of the weights: a wrong level weight, a dropped
  level (the (3, 3) products are level 6) or a wrong sign fails it;
* random operands agree with FP64 to the 35-bit representation (a 28-bit, four-digit version
  misses this bound by two orders of magnitude);
* the chunking does not change the exact case;
* screening: the function order only moves F by the rounding of the transposed elements, and
  the blocks left out at threshold 2^-S move each element by at most nchunk 2^-S, while a
  threshold that the bound does not cover (the exponents taken one too small) breaks it."""
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


def _local_system():
    # three separated centres, each with 16 compact and 8 diffuse functions (s, x, y, z times
    # four and two exponents): the chunk blocks of one centre's compact functions are empty
    # at the other centres' points
    rng = np.random.default_rng(9)
    centres = np.array([[0.0, 0.0, 0.0], [3.5, 0.0, 0.0], [0.0, 4.0, 1.0]])
    pts = np.concatenate([c + rng.normal(size=(640, 3)) * 0.6 for c in centres])
    wbw = rng.uniform(0.01, 0.2, size=len(pts))
    cols = []
    for c in centres:
        r2 = ((pts - c) ** 2).sum(1)
        for a in (60.0, 30.0, 15.0, 8.0, 0.8, 0.3):
            g = np.exp(-a * r2)
            cols += [g, (pts[:, 0] - c[0]) * g, (pts[:, 1] - c[1]) * g, (pts[:, 2] - c[2]) * g]
    phi = np.stack(cols, 1)              # [1920, 72]: 5 blocks, the last one partial
    v = 1.0 / (0.3 + ((pts - centres[0]) ** 2).sum(1)) + 0.5
    return phi, wbw, v


def test_int8_screening_order_and_bound():
    phi, wbw, v = _local_system()
    chunk = 128
    nchunk = -(-phi.shape[0] // chunk)
    order = core.int8_compact_order(phi, wbw, chunk)
    assert sorted(order) == list(range(phi.shape[1]))
    F_id = core.int8_fock(phi, wbw, v, chunk)
    F0 = core.int8_fock(phi, wbw, v, chunk, order=order)
    assert np.abs(F0 - F_id).max() <= 1e-9 * np.abs(F_id).max()
    for S in (2, 12, 30):
        st = {}
        FS = core.int8_fock(phi, wbw, v, chunk, screen=S, order=order, stats=st)
        d = np.abs(FS - F0).max()
        print('screen 2^-%d: %d of %d chunk blocks left out, max|dF| %.3e <= %.3e' % (
            S, st['skipped'], st['kept'] + st['skipped'], d, nchunk * 2.0 ** -S))
        assert st['skipped'] > 0 and d > 0.0
        assert d <= nchunk * 2.0 ** -S * (1 + 1e-6) + 1e-15 * np.abs(F0).max()
        assert np.array_equal(FS, FS.T)
    # a planted mistake: the bound without the 2^lg points per chunk leaves out blocks it does
    # not cover, and the element bound fails
    real_ceil = np.ceil
    try:
        core.np.ceil = lambda x: real_ceil(x) - 7.0
        FW = core.int8_fock(phi, wbw, v, chunk, screen=2, order=order)
    finally:
        core.np.ceil = real_ceil
    assert np.abs(FW - F0).max() > nchunk * 2.0 ** -2
