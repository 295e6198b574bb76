"""ORACLE (test infrastructure only): independent M-scheme shell-model Hamiltonian.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import anything under ``oracle/``; the product never does.

This is a textbook second-quantised shell-model code that shares nothing with
the J-scheme path: proton and neutron Slater determinants, uncoupled
antisymmetrised matrix elements built from the isospin-formalism JT TBMEs,

    <ab|V|cd>_T=1 (pp, nn) = sum_J CG CG sqrt((1+d_ab)(1+d_cd)) V_J,T=1(ab,cd)
    <a_p b_n|V|c_p d_n>     = sum_J CG CG sqrt((1+d_ab)(1+d_cd))/2 (V_J1 + V_J0),

and H = sum eps n + sum_{a<b,c<d} V a+a+aa (pp, nn) + sum V_pn a+ b+ b a.
Its M=M0 spectrum is the union over J >= |M0| of the J-scheme spectra, which
pins the pn J-scheme H*v (NuTRL ``multiplyx/a/b``,
``src/NuTRL/NuCoreX.f90:76-226``) together with the generated operator files
without using either.  Clebsch-Gordan coefficients come from sympy.
"""
from __future__ import annotations

from functools import lru_cache
from itertools import combinations
from typing import Dict, List, Sequence, Tuple

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


@lru_cache(maxsize=None)
def _cg(j1: int, j2: int, J: int, m1: int, m2: int, M: int) -> float:
    """<j1/2 m1/2 j2/2 m2/2 | J/2 M/2> via sympy (doubled arguments)."""
    from sympy import Rational
    from sympy.physics.wigner import clebsch_gordan
    if m1 + m2 != M or abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return 0.0
    if not (abs(j1 - j2) <= J <= j1 + j2):
        return 0.0
    h = Rational(1, 2)
    return float(clebsch_gordan(j1 * h, j2 * h, J * h, m1 * h, m2 * h, M * h))


def read_int_file(path: str):
    """SPE line then 'a b c d J T V' (1-based orbits, undoubled J, T)."""
    with open(path) as f:
        head = f.readline().split()
        spe = [float(x) for x in head[1:]]
        tb = {}
        for line in f:
            p = line.split()
            if len(p) >= 7:
                a, b, c, d, J, T = (int(x) - (1 if i < 4 else 0) for i, x in enumerate(p[:6]))
                tb[(a, b, c, d, J, T)] = float(p[6])
    return spe, tb


class MSchemeOracle:
    def __init__(self, j2s: Sequence[int], spe: Sequence[float], tbme: Dict, nprot: int, nneut: int):
        self.j2s = list(j2s)
        self.spe = list(spe)
        self.tb = tbme
        self.np_, self.nn_ = nprot, nneut
        self.sps = [(o, m) for o, j2 in enumerate(self.j2s) for m in range(-j2, j2 + 1, 2)]

    # JT matrix element for any orbit order; J, T undoubled
    def vjt(self, a, b, c, d, J, T) -> float:
        ph = 1.0
        ja, jb, jc, jd = (self.j2s[x] for x in (a, b, c, d))
        if a > b:
            ph *= (-1) ** (((ja + jb) // 2 - J - T) % 2)
            a, b = b, a
        if c > d:
            ph *= (-1) ** (((jc + jd) // 2 - J - T) % 2)
            c, d = d, c
        if (a, b) > (c, d):
            a, b, c, d = c, d, a, b
        return ph * self.tb.get((a, b, c, d, J, T), 0.0)

    def v_like(self, s1, s2, s3, s4) -> float:
        (a, ma), (b, mb), (c, mc), (d, md) = (self.sps[s] for s in (s1, s2, s3, s4))
        M = ma + mb
        if mc + md != M:
            return 0.0
        n = np.sqrt((2.0 if a == b else 1.0) * (2.0 if c == d else 1.0))
        tot = 0.0
        for J2 in range(0, 2 * max(self.j2s) + 1, 2):
            c1 = _cg(self.j2s[a], self.j2s[b], J2, ma, mb, M)
            c2 = _cg(self.j2s[c], self.j2s[d], J2, mc, md, M)
            if c1 and c2:
                tot += c1 * c2 * n * self.vjt(a, b, c, d, J2 // 2, 1)
        return tot

    def v_pn(self, sp1, sn2, sp3, sn4) -> float:
        (a, ma), (b, mb), (c, mc), (d, md) = (self.sps[s] for s in (sp1, sn2, sp3, sn4))
        M = ma + mb
        if mc + md != M:
            return 0.0
        n = np.sqrt((2.0 if a == b else 1.0) * (2.0 if c == d else 1.0)) / 2.0
        tot = 0.0
        for J2 in range(0, 2 * max(self.j2s) + 1, 2):
            c1 = _cg(self.j2s[a], self.j2s[b], J2, ma, mb, M)
            c2 = _cg(self.j2s[c], self.j2s[d], J2, mc, md, M)
            if c1 and c2:
                tot += c1 * c2 * n * (self.vjt(a, b, c, d, J2 // 2, 1) + self.vjt(a, b, c, d, J2 // 2, 0))
        return tot

    def _dets(self, n: int):
        out = {}
        for c in combinations(range(len(self.sps)), n):
            M = sum(self.sps[i][1] for i in c)
            out.setdefault(M, []).append(c)
        return out

    def hamiltonian(self, M2: int) -> sp.csr_matrix:
        S = len(self.sps)
        pd = self._dets(self.np_)
        nd = self._dets(self.nn_)
        basis: List[Tuple[Tuple[int, ...], Tuple[int, ...]]] = []
        for Mp, plist in sorted(pd.items()):
            for q in plist:
                for r in nd.get(M2 - Mp, []):
                    basis.append((q, r))
        index = {b: i for i, b in enumerate(basis)}
        pairs = [(a, b) for a in range(S) for b in range(a + 1, S)]
        vlike = {}
        for (a, b) in pairs:
            for (c, d) in pairs:
                v = self.v_like(a, b, c, d)
                if abs(v) > 0:
                    vlike[(a, b, c, d)] = v
        vpn = {}
        for a in range(S):
            for b in range(S):
                for c in range(S):
                    for d in range(S):
                        v = self.v_pn(a, b, c, d)
                        if abs(v) > 0:
                            vpn[(a, b, c, d)] = v
        rows, cols, vals = [], [], []

        def ann(det, s):
            if s not in det:
                return None, 0
            i = det.index(s)
            return det[:i] + det[i + 1:], i

        def cre(det, s):
            if s in det:
                return None, 0
            i = sum(1 for x in det if x < s)
            return det[:i] + (s,) + det[i:], i

        def two_body(det, V):
            out = []
            for (a, b, c, d), v in V.items():
                d1, s1 = ann(det, c)
                if d1 is None:
                    continue
                d2, s2 = ann(d1, d)
                if d2 is None:
                    continue
                d3, s3 = cre(d2, b)
                if d3 is None:
                    continue
                d4, s4 = cre(d3, a)
                if d4 is None:
                    continue
                out.append((d4, v * (-1) ** (s1 + s2 + s3 + s4)))
            return out

        for col, (q, r) in enumerate(basis):
            diag = sum(self.spe[self.sps[s][0]] for s in q) + sum(self.spe[self.sps[s][0]] for s in r)
            rows.append(col); cols.append(col); vals.append(diag)
            for q2, v in two_body(q, vlike):
                rows.append(index[(q2, r)]); cols.append(col); vals.append(v)
            for r2, v in two_body(r, vlike):
                rows.append(index[(q, r2)]); cols.append(col); vals.append(v)
            # pn: a+_a(p) b+_b(n) b_d(n) a_c(p): proton and neutron operators commute
            # in this product-space representation (protons ordered before neutrons)
            for (a, b, c, d), v in vpn.items():
                q1, s1 = ann(q, c)
                if q1 is None:
                    continue
                r1, s2 = ann(r, d)
                if r1 is None:
                    continue
                r2, s3 = cre(r1, b)
                if r2 is None:
                    continue
                q2, s4 = cre(q1, a)
                if q2 is None:
                    continue
                rows.append(index[(q2, r2)]); cols.append(col)
                vals.append(v * (-1) ** (s1 + s2 + s3 + s4))
        n = len(basis)
        return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))

    def lowest(self, M2: int, k: int) -> np.ndarray:
        H = self.hamiltonian(M2)
        if H.shape[0] <= 3000:
            return np.linalg.eigvalsh(H.toarray())[:k]
        w = spla.eigsh(H, k=k, which="SA", tol=1e-12)[0]
        return np.sort(w)
