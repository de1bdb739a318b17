"""The random draws of the reference trainer, restated bit for bit.

The Taylor augmentation of src/nnp/nnptefr2.cpp:2383-2482 draws from a
std::mt19937 (seeded as :2160-2168 do: rngen = mt19937(seed_global), the
first output as the rank's seed, rngen = mt19937(that seed)) through the
reference's own distribution classes (src/rand/rng.cpp:64-147) and
std::uniform_real_distribution<double>(-1, 1).  This module restates:

* MT19937 (Matsumoto and Nishimura 1998): init_genrand seeding, the twist in
  blocks of 624 words, tempering -- the engine std::mt19937 specifies;
* generate_canonical<double, 53> over it as libstdc++ computes it (two
  32-bit outputs, sum = u0 + u1 2^32 in double, divided by 2^64, clamped below
  1) and uniform_real_distribution(a, b) = canonical * (b - a) + a;
* the reference's Uniform, Laplace, Normal (Box-Muller pairs with the second
  value cached), Logistic, Sech and Cosine samplers, with u = gen() / 2^32-1.

Scalar arithmetic goes through the platform libm, as the C++ build does, so
that every value matches it on the same host: Python's math module for log,
sqrt and tan, and libm's sincos for a sine and cosine of one angle, which GCC
emits in place of separate sin and cos calls (their last bits can differ).
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math

import numpy as np

_N, _M = 624, 397
_MATRIX_A = 0x9908B0DF
_UPPER, _LOWER = 0x80000000, 0x7FFFFFFF
_MAX = 4294967295.0                       # std::mt19937::max()

_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
_libm.sincos.restype = None


def sincos(x: float) -> tuple:
    """(sin x, cos x) from libm's sincos."""
    s, c = ctypes.c_double(), ctypes.c_double()
    _libm.sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


class MT19937:
    """std::mt19937 with its default (init_genrand) seeding."""

    def __init__(self, seed: int = 5489):
        mt = [0] * _N
        mt[0] = seed & 0xFFFFFFFF
        for i in range(1, _N):
            mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self._mt = np.array(mt, dtype=np.uint64)
        self._out = []
        self._pos = 0

    def _twist(self):
        mt = self._mt
        upper, lower, mag = np.uint64(_UPPER), np.uint64(_LOWER), np.uint64(_MATRIX_A)

        def block(i0, i1, j0):
            # mt[i] for i in [i0, i1) from mt[i], mt[i + 1] and mt[i + 397 mod 624]
            y = (mt[i0:i1] & upper) | (mt[i0 + 1:i1 + 1] & lower)
            mt[i0:i1] = mt[j0:j0 + (i1 - i0)] ^ (y >> np.uint64(1)) ^ np.where(y & np.uint64(1), mag, np.uint64(0))

        block(0, _N - _M, _M)                          # partners not yet rewritten
        for start in range(_N - _M, _N - 1, _N - _M):  # partners rewritten in the block before
            stop = min(start + (_N - _M), _N - 1)
            block(start, stop, start + _M - _N)
        y = (mt[_N - 1] & upper) | (mt[0] & lower)
        mt[_N - 1] = mt[_M - 1] ^ (y >> np.uint64(1)) ^ (mag if int(y) & 1 else np.uint64(0))
        # tempering
        y = mt.copy()
        y ^= y >> np.uint64(11)
        y ^= (y << np.uint64(7)) & np.uint64(0x9D2C5680)
        y ^= (y << np.uint64(15)) & np.uint64(0xEFC60000)
        y ^= y >> np.uint64(18)
        self._out = [int(v) & 0xFFFFFFFF for v in y]
        self._pos = 0

    def __call__(self) -> int:
        if self._pos >= len(self._out):
            self._twist()
        v = self._out[self._pos]
        self._pos += 1
        return v


def canonical(gen: MT19937) -> float:
    """generate_canonical<double, 53>(gen) as libstdc++ computes it."""
    s = float(gen())
    s += float(gen()) * 4294967296.0
    r = s / 18446744073709551616.0
    return r if r < 1.0 else math.nextafter(1.0, 0.0)


def uniform_real(gen: MT19937, a: float, b: float) -> float:
    """std::uniform_real_distribution<double>(a, b)(gen)."""
    return canonical(gen) * (b - a) + a


def _unit(gen: MT19937) -> float:
    return float(gen()) / _MAX


def _unit_nonzero(gen: MT19937) -> float:
    p = 0.0
    while p == 0.0:
        p = _unit(gen)
    return p


class Distribution:
    """rng::dist::Uniform / Laplace / Normal / Logistic / Sech / Cosine with
    location 0 and scale ``s`` (src/rand/rng.cpp:64-147)."""

    NAMES = ("uniform", "laplace", "normal", "logistic", "sech", "cosine")

    def __init__(self, name: str, s: float):
        if name not in self.NAMES:
            raise ValueError(f"unknown distribution {name!r} (one of {', '.join(self.NAMES)})")
        self.name, self.s = name, float(s)
        self._count, self._y = 0, 0.0

    def rand(self, gen: MT19937) -> float:
        s = self.s
        if self.name == "uniform":                     # U(0, s)
            return _unit(gen) * (s - 0.0) + 0.0
        if self.name == "laplace":
            p = _unit_nonzero(gen) - 0.5
            sgn = (p > 0.0) - (p < 0.0)
            return 0.0 - 1.0 * s * sgn * math.log(1.0 - 2.0 * abs(p))
        if self.name == "normal":                      # Box-Muller, the second value kept
            if self._count % 2 == 0:
                r1 = _unit_nonzero(gen)
                r2 = _unit(gen)
                mag = s * math.sqrt(-2.0 * math.log(r1))
                sn, cs = sincos(2.0 * math.pi * r2)
                x = mag * cs
                self._y = mag * sn
                self._count += 1
                return x
            self._count += 1
            return self._y
        if self.name == "logistic":
            p = _unit_nonzero(gen)
            return 0.0 + s * math.log(p / (1.0 - p))
        if self.name == "sech":
            p = _unit_nonzero(gen)
            return 0.0 + s * 2.0 / math.pi * math.log(math.tan(0.5 * math.pi * p))
        p = _unit_nonzero(gen)                         # cosine
        if p < 0.5:
            t = math.sqrt(-2.0 * math.log(p))
            return 0.0 + s * (0.820251 - 0.806770 * t + 0.093247 * t * t)
        t = math.sqrt(-2.0 * math.log(1.0 - p))
        return 0.0 - s * (0.820251 - 0.806770 * t + 0.093247 * t * t)


def trainer_generator(seed_global: int) -> MT19937:
    """The generator of the reference trainer on one rank: mt19937(seed_global),
    its first output (stored as an int) seeds the rank's mt19937."""
    first = MT19937(seed_global)()
    return MT19937(first)
