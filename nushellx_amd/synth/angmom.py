"""Angular-momentum coefficients for the operator-file generator.

All arguments are DOUBLED (2j, 2m), as everywhere in NuShellX.  Values are
computed from the Racah closed forms with exact integer arithmetic and rounded
once to float64.

* ``cg(j1, j2, j3, m1, m2, m3)`` = <j1 m1 j2 m2 | j3 m3>, the argument order of
  NuShellX's ``clebg`` (``src/NuCommon/ICommon.FI:1083``).
* ``sixj(a, b, c, d, e, f)`` = {a b c; d e f}.
* ``racah(j1, j2, j3, j4, J, K)`` = W(j1 j2 j3 j4; J K)
  = (-1)^{(j1+j2+j3+j4)/2} {j1 j2 J; j4 j3 K}, the definition of
  ``src/NuCommon/NuCommon.f90:3255-3265``.
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from math import factorial, sqrt


def _tri(a: int, b: int, c: int) -> bool:
    return (a + b + c) % 2 == 0 and abs(a - b) <= c <= a + b


@lru_cache(maxsize=None)
def cg(j1: int, j2: int, j3: int, m1: int, m2: int, m3: int) -> float:
    if m1 + m2 != m3:
        return 0.0
    if not _tri(j1, j2, j3):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    if (j1 + m1) % 2 or (j2 + m2) % 2 or (j3 + m3) % 2:
        return 0.0
    # integer (undoubled) combinations
    a = (j1 + j2 - j3) // 2
    b = (j1 - j2 + j3) // 2
    c = (-j1 + j2 + j3) // 2
    d = (j1 + j2 + j3) // 2 + 1
    pref = Fraction((j3 + 1) * factorial(a) * factorial(b) * factorial(c), factorial(d))
    pref *= (factorial((j1 + m1) // 2) * factorial((j1 - m1) // 2)
             * factorial((j2 + m2) // 2) * factorial((j2 - m2) // 2)
             * factorial((j3 + m3) // 2) * factorial((j3 - m3) // 2))
    s = Fraction(0)
    kmin = max(0, (j2 - j3 - m1) // 2, (j1 - j3 + m2) // 2)
    kmax = min(a, (j1 - m1) // 2, (j2 + m2) // 2)
    for k in range(kmin, kmax + 1):
        den = (factorial(k) * factorial(a - k) * factorial((j1 - m1) // 2 - k)
               * factorial((j2 + m2) // 2 - k) * factorial((j3 - j2 + m1) // 2 + k)
               * factorial((j3 - j1 - m2) // 2 + k))
        s += Fraction((-1) ** k, den)
    val = sqrt(pref) * float(s)
    return float(val)


def _delta(a: int, b: int, c: int) -> Fraction:
    return Fraction(factorial((a + b - c) // 2) * factorial((a - b + c) // 2)
                    * factorial((-a + b + c) // 2), factorial((a + b + c) // 2 + 1))


@lru_cache(maxsize=None)
def sixj(a: int, b: int, c: int, d: int, e: int, f: int) -> float:
    if not (_tri(a, b, c) and _tri(a, e, f) and _tri(d, b, f) and _tri(d, e, c)):
        return 0.0
    pref2 = _delta(a, b, c) * _delta(a, e, f) * _delta(d, b, f) * _delta(d, e, c)
    t1 = (a + b + c) // 2
    t2 = (a + e + f) // 2
    t3 = (d + b + f) // 2
    t4 = (d + e + c) // 2
    p1 = (a + b + d + e) // 2
    p2 = (a + c + d + f) // 2
    p3 = (b + c + e + f) // 2
    s = Fraction(0)
    for t in range(max(t1, t2, t3, t4), min(p1, p2, p3) + 1):
        den = (factorial(t - t1) * factorial(t - t2) * factorial(t - t3) * factorial(t - t4)
               * factorial(p1 - t) * factorial(p2 - t) * factorial(p3 - t))
        s += Fraction((-1) ** t * factorial(t + 1), den)
    return float(sqrt(pref2) * float(s)) if s != 0 else 0.0


def racah(j1: int, j2: int, j3: int, j4: int, J: int, K: int) -> float:
    ph = -1.0 if ((j1 + j2 + j3 + j4) // 2) % 2 else 1.0
    return ph * sixj(j1, j2, J, j4, j3, K)
