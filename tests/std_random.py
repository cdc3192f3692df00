"""The C++ standard library's std::mt19937 and libstdc++'s
std::uniform_real_distribution<double>, restated in Python so that a test
can rebuild the right-hand side drawn by the reference's util.hpp:50-54
rand(x) for a known seed.

MT19937: Matsumoto and Nishimura (1998), 32-bit, state size 624, the
parameters of [rand.predef] in the C++ standard; seeding by the standard's
linear recurrence (init_genrand).  uniform_real_distribution<double>(a, b):
libstdc++ draws generate_canonical<double, 53> from two 32-bit outputs,
u = (g1 + g2 * 2^32) / 2^64 in double arithmetic (clamped below 1), and
returns a + (b - a) * u.
"""
import numpy as np

_N, _M = 624, 397
_MATRIX_A, _UPPER, _LOWER = 0x9908B0DF, 0x80000000, 0x7FFFFFFF


class MT19937:
    def __init__(self, seed=5489):
        mt = [0] * _N
        mt[0] = seed & 0xFFFFFFFF
        for i in range(1, _N):
            mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.mt = mt
        self.i = _N

    def _twist(self):
        mt = self.mt
        for k in range(_N):
            y = (mt[k] & _UPPER) | (mt[(k + 1) % _N] & _LOWER)
            mt[k] = mt[(k + _M) % _N] ^ (y >> 1) ^ (_MATRIX_A if y & 1 else 0)
        self.i = 0

    def __call__(self):
        if self.i >= _N:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def uniform_real(gen, count, a=0.0, b=1.0):
    """count draws of std::uniform_real_distribution<double>(a, b)(gen)."""
    out = np.empty(count)
    one_below = np.nextafter(1.0, 0.0)
    for i in range(count):
        s = float(gen()) + float(gen()) * 4294967296.0
        u = s / 18446744073709551616.0
        if u >= 1.0:
            u = one_below
        out[i] = (b - a) * u + a
    return out


def reference_rand(count, seed):
    """util.hpp:50-54 rand(x) with std::random_device returning `seed`."""
    return uniform_real(MT19937(seed), count)
