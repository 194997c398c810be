"""64-bit Mersenne Twister (MT19937-64, Matsumoto & Nishimura 2004) and the
documented bit -> double map used for the synthetic benchmark inputs, so that
the GPU runs, the oracle and the golden generator draw identical inputs without
depending on a library's distribution code.

    u = (x >> 11) * 2**-53        # 53 random bits -> [0, 1)

The generator's output sequence is the one C++ std::mt19937_64 produces
(its 10000th output for the default seed 5489 is 9981545732273789042; checked
in tests/test_oracle.py).
"""
import numpy as np

_N, _M = 312, 156
_A = 0xB5026F5AA96619E9
_UM = 0xFFFFFFFF80000000
_LM = 0x7FFFFFFF
_MASK = 0xFFFFFFFFFFFFFFFF


class MT19937_64:
    def __init__(self, seed=5489):
        mt = [0] * _N
        mt[0] = seed & _MASK
        for i in range(1, _N):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & _MASK
        self.mt = mt
        self.i = _N

    def _twist(self):
        mt = self.mt
        for i in range(_N):
            x = (mt[i] & _UM) | (mt[(i + 1) % _N] & _LM)
            xa = x >> 1
            if x & 1:
                xa ^= _A
            mt[i] = mt[(i + _M) % _N] ^ xa
        self.i = 0

    def next(self):
        if self.i >= _N:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK


def uniform(seed, n, lo=0.0, hi=1.0):
    """n doubles lo + (hi - lo) * u with u = (x >> 11) * 2**-53."""
    g = MT19937_64(seed)
    u = np.array([(g.next() >> 11) for _ in range(n)], dtype=np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def cfg1_input():
    """The 128x128x1 binary field of the benchmark's CPU-runnable case:
    phi_i = U[-0.1, 0.1] i.i.d. from mt19937_64 seeded 20240."""
    return uniform(20240, 128 * 128, -0.1, 0.1)


def minima_input(case):
    """Initial state of a ref_minima case, regenerated from its recipe
    (mt19937_64 draws, this module) so the fixture stores no inputs.

    stripes: the 128x128 benchmark field (phi ~ U[-0.1, 0.1], seed 20240)
      plus +-bias on the two halves of x, which sets two interfaces.
    slabs: nslab slabs of pure fluids along x, each site mixed with
      amplitude * U[-1, 1] noise per fluid, renormalised to sum 1."""
    grid, nf = case["grid"], case["nfluid"]
    ns = int(np.prod(grid))
    if case["init"] == "stripes":
        x0 = cfg1_input() if tuple(grid) == (128, 128, 1) else uniform(case["seed"], ns, -0.1, 0.1)
        xg = np.arange(grid[0])
        bias = np.where(np.sin(2 * np.pi * (xg + 0.5) / grid[0]) > 0, case["amplitude"], -case["amplitude"])
        return x0 + np.repeat(bias, grid[1] * grid[2])
    u = uniform(case["seed"], ns * nf, -1.0, 1.0).reshape(ns, nf)
    lab = np.arange(grid[0]) * case["nslab"] // grid[0]
    c = np.zeros((grid[0], nf))
    c[np.arange(grid[0]), lab] = 1.0
    c = np.repeat(c[:, None, :], grid[1] * grid[2], axis=1).reshape(-1, nf)
    x0 = c + case["amplitude"] * u
    return (x0 / x0.sum(axis=1, keepdims=True)).reshape(-1)
