"""Writes tests/golden/angmom.json: exact 6j and Clebsch-Gordan values from sympy.

These pin the recoupling coefficients used by the oracle (oracle/refcpu.c,
binomial form of src/NuCommon/NuCommon.f90:3268-3309), the product
(nushellx_amd/csrc/nsx_model.cpp, Racah single sum) and the generator
(nushellx_amd/synth/angmom.py).  Arguments are doubled, as in NuShellX.
Run: python tests/golden/make_golden.py
"""
import itertools
import json
import os

from sympy import Rational
from sympy.physics.wigner import clebsch_gordan, wigner_6j

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "angmom.json")


def main():
    h = Rational(1, 2)
    sixj = []
    # doubled arguments up to 2*(7/2 + 7/2) as met in sd/pf recouplings: sample deterministically
    vals = list(range(0, 15))
    for a, b, c, d, e, f in itertools.product(vals, repeat=6):
        if (a + b + c) % 2 or (c + d + e) % 2 or (a + e + f) % 2 or (b + d + f) % 2:
            continue
        if not (abs(a - b) <= c <= a + b and abs(c - d) <= e <= c + d and abs(a - e) <= f <= a + e
                and abs(b - d) <= f <= b + d):
            continue
        if (a * 7 + b * 5 + c * 3 + d * 11 + e * 13 + f * 17) % 97 != 0:
            continue  # a deterministic ~1% subsample
        sixj.append([a, b, c, d, e, f, float(wigner_6j(a * h, b * h, c * h, d * h, e * h, f * h))])
    # large-J cases as met by side J up to 2J = 30 (pf 6 particles)
    for (a, b, c, d, e, f) in [(30, 24, 6, 28, 26, 8), (24, 30, 10, 22, 28, 14), (20, 20, 0, 14, 14, 14),
                               (29, 27, 4, 25, 23, 6), (16, 18, 2, 12, 14, 8)]:
        sixj.append([a, b, c, d, e, f, float(wigner_6j(a * h, b * h, c * h, d * h, e * h, f * h))])
    cg = []
    for j1 in range(0, 8):
        for j2 in range(0, 8):
            for j3 in range(abs(j1 - j2), j1 + j2 + 1, 2):
                for m1 in range(-j1, j1 + 1, 2):
                    for m2 in range(-j2, j2 + 1, 2):
                        m3 = m1 + m2
                        if abs(m3) > j3 or (j1 + j2 + j3 + m1) % 3:
                            continue
                        cg.append([j1, j2, j3, m1, m2, m3,
                                   float(clebsch_gordan(j1 * h, j2 * h, j3 * h, m1 * h, m2 * h, m3 * h))])
    with open(OUT, "w") as fh:
        json.dump({"sixj": sixj, "cg": cg}, fh)
    print(len(sixj), "6j values,", len(cg), "CG values ->", OUT)


if __name__ == "__main__":
    main()
