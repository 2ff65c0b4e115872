"""std::mt19937 and libstdc++'s std::uniform_real_distribution<double>, restated in Python.

MT19937 is Matsumoto and Nishimura's generator (ACM TOMACS 8 (1998) 3-30) with the
32-bit seeding of the C++ standard ([rand.eng.mers]: x_i = 1812433253 (x_{i-1} xor
(x_{i-1} >> 30)) + i).  libstdc++ draws a double in [0, 1) by generate_canonical<double, 53>:
two 32-bit outputs g1, g2 combined as (g1 + g2 * 2^32) / 2^64 in double arithmetic (the
result is moved below 1 if it rounds up to 1), and uniform_real_distribution(a, b) returns
canonical * (b - a) + a.  PdfBase::fillMCDataSimple draws its events this way
(src/goofit/PdfBase.cpp:172-199), so these give the events a seed must produce.
"""
import math

_N, _M = 624, 397


class MT19937:
    def __init__(self, seed=5489):
        mt = [seed & 0xFFFFFFFF]
        for i in range(1, _N):
            prev = mt[-1]
            mt.append((1812433253 * (prev ^ (prev >> 30)) + i) & 0xFFFFFFFF)
        self.mt, self.i = mt, _N

    def _twist(self):
        mt = self.mt
        for k in range(_N):
            y = (mt[k] & 0x80000000) | (mt[(k + 1) % _N] & 0x7FFFFFFF)
            mt[k] = mt[(k + _M) % _N] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
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
        return y


def canonical(gen):
    """generate_canonical<double, 53>(gen) for a 32-bit generator."""
    s = float(gen())
    s += float(gen()) * 4294967296.0
    r = s / 18446744073709551616.0
    return r if r < 1.0 else math.nextafter(1.0, 0.0)


def uniform(gen, a, b):
    """std::uniform_real_distribution<double>(a, b)(gen)."""
    return canonical(gen) * (b - a) + a
