"""Separable pn operators (``.dop``) as program NuDiop writes them.

NuDiop (``src/NuDiop/NuDiop.f90``) reads the pn multipole interaction of an
``.oph`` file (one record per key (p p' n n' 2lam) with the pair parities), and
for every (2lam, parity) forms the matrix ``V(P_i, conj(P_j))`` over the
proton pairs ``P`` and the neutron pairs ``conj(P)`` on the same orbits
(``:48-159``).  It diagonalises that matrix (``:186-188``) and writes, for each
eigenvalue whose ``e / sqrt(2lam+1)`` lies outside the cut, the eigenvalue and
the eigenvector components above ``mcut`` as lines
``p p' n n' amplitude`` (``:204-233``).  With no cut the terms reproduce the
interaction exactly: ``V(P_i, conj(P_j)) = sum_n e_n u_n(i) u_n(j)``.

NuTRLZ and NuXpct read the file back as a sum of separable terms
(``src/NuTRLZ/NuCoreXZ.f90:1613-1887``).
"""
from __future__ import annotations

import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .fortio import read_records

Pair = Tuple[int, int]
Line = Tuple[int, int, int, int, float]
Section = Tuple[int, int, List[Tuple[float, List[Line]]]]


def read_oph(path: str):
    """Records of an ``.oph`` file: [(key bytes, meV, pl, pr)], max_lmo (``src/NuOp/NuOp.f90:410-418``)."""
    recs = read_records(path)
    head = next(recs)
    n, max_lmo = np.frombuffer(head[:8], dtype="<i4")
    items = []
    for _ in range(int(n)):
        r = next(recs)
        me = float(np.frombuffer(r[8:16], dtype="<f8")[0])
        pl, pr = (int(x) for x in np.frombuffer(r[16:24], dtype="<i4"))
        items.append((bytes(r[:8]), me, pl, pr))
    return items, int(max_lmo)


def _distinct(pairs: Sequence[Pair]) -> List[Pair]:
    out: List[Pair] = []
    for p in pairs:
        if p not in out:
            out.append(p)
    return out


def _code(p: Pair) -> int:
    return 100 * p[0] + p[1]


def diop_sections(items, max_lmo: int, cut: float = 0.0, mcut: float = 0.0,
                  cut_table: Optional[Dict[int, Tuple[float, float]]] = None) -> List[Section]:
    """The (2lam, parity) sections NuDiop writes.  ``cut``/``mcut`` are its third and
    fourth arguments; ``cut_table`` maps a multipole index to (lower, upper) cuts
    as its optional ``diop.cut`` file does (index 0 keeps every nonzero term)."""
    if cut < 0:
        raise ValueError("cut must be >= 0")
    chl = _distinct([(k[0], k[1]) for k, _, _, _ in items])
    chr_ = _distinct([(k[2], k[3]) for k, _, _, _ in items])
    if len(chl) != len(chr_):
        raise ValueError("cr /= cl: pair lists not symmetric")
    shl = [s for p in chl for s in p]
    shr = [s for p in chr_ for s in p]
    if max(shr) - min(shr) != max(shl) - min(shl):
        raise ValueError("shells not symmetric")
    half = max(shr) // 2
    # the right pair on the same orbits as each left pair (:101-114)
    conj: List[Pair] = []
    for p in chl:
        m = None
        for q in chr_:
            if (p[0] + half, p[1] + half) == q or (p[0] - half, p[1] - half) == q:
                m = q
        if m is None:
            raise ValueError(f"no conjugate pair for {p}")
        conj.append(m)
    n = len(chl)
    cutp = np.zeros(18)
    cutn = np.zeros(18)
    cutp[1:] = cut
    cutn[1:] = -cut
    for k, (lo, hi) in (cut_table or {}).items():
        cutn[k], cutp[k] = lo, hi
    left_of = {_code(p): i for i, p in enumerate(chl)}
    right_of = {_code(p): i for i, p in enumerate(conj)}
    out: List[Section] = []
    for lm in range(0, max_lmo + 1, 2):
        for pt in (-1, 1):
            mat = np.zeros((n, n))
            nonz = 0
            for key, me, pl, pr in items:
                if key[4] != lm or pl != pr or pl != pt:
                    continue
                il, ir = _code((key[0], key[1])), _code((key[2], key[3]))
                if il > ir:
                    continue
                if ir not in right_of:
                    raise ValueError(f"right pair {key[2], key[3]} has no left partner")
                mat[left_of[il], right_of[ir]] += me
                nonz += 1
            if nonz == 0:
                out.append((lm, pt, []))
                continue
            mapr = [i for i in range(n) if np.abs(mat[:, i]).sum() != 0.0]
            mapl = [i for i in range(n) if np.abs(mat[i, :]).sum() != 0.0]
            if len(mapr) != len(mapl):
                raise ValueError("submatrix not square")
            if not mapl:
                out.append((lm, pt, []))
                continue
            sub = mat[np.ix_(mapl, mapr)]
            alpha, vec = np.linalg.eigh(sub, UPLO="L")  # ascending, as tqli + indexx
            rt = np.sqrt(lm + 1.0)
            terms = []
            for i in range(len(alpha)):
                a = alpha[i] / rt
                if not ((a > cutp[lm // 2] and a > 0.0) or (a < cutn[lm // 2] and a < 0.0)):
                    continue
                lines: List[Line] = []
                for j in range(len(mapl)):
                    u = vec[j, i]
                    pl_, pr_ = chl[mapl[j]], conj[mapr[j]]
                    if abs(u) > mcut and _code(pl_) < _code(pr_):
                        lines.append((pl_[0], pl_[1], pr_[0], pr_[1], float(u)))
                terms.append((float(alpha[i]), lines))
            out.append((lm, pt, terms))
    return out


def write_dop(path: str, sections: Sequence[Section], max_lmo: int, exact: bool = False) -> None:
    """NuDiop's text layout (``(i4)``, ``(2i4)``, ``(f11.7)``, ``(i6)``, ``(4i4,f11.7)``);
    ``exact`` writes the reals with 17 significant digits instead (the reader is
    list-directed, so both parse the same way)."""
    fr = (lambda x: f" {x:.17e}") if exact else (lambda x: f"{x:11.7f}")
    with open(path, "w") as f:
        f.write(f"{max_lmo:4d}\n")
        for lm, pt, terms in sections:
            f.write(f"{lm:4d}{pt:4d}\n")
            f.write(f"{len(terms):4d}\n")
            for e, lines in terms:
                f.write(fr(e) + "\n")
                f.write(f"{len(lines):6d}\n")
                for s1, s2, s3, s4, u in lines:
                    f.write(f"{s1:4d}{s2:4d}{s3:4d}{s4:4d}" + fr(u) + "\n")


def write_model_dop(intfile: str, out_prefix: Optional[str] = None, cut: float = 0.0, mcut: float = 0.0,
                    exact: bool = False) -> str:
    """``<out_prefix>.dop`` from ``<intfile>.oph``; returns the prefix (NuTRLZ's 5th argument)."""
    items, max_lmo = read_oph(intfile + ".oph")
    prefix = out_prefix or intfile
    write_dop(prefix + ".dop", diop_sections(items, max_lmo, cut, mcut), max_lmo, exact)
    return prefix


def separable_matrix(sections: Sequence[Section], lm: int, pt: int) -> Dict[Tuple[Pair, Pair], float]:
    """sum_n e_n u_n(P) u_n(N) over the terms of one section, keyed by (P, N)."""
    out: Dict[Tuple[Pair, Pair], float] = {}
    for l, p, terms in sections:
        if (l, p) != (lm, pt):
            continue
        for e, lines in terms:
            for a in lines:
                for b in lines:
                    k = ((a[0], a[1]), (b[2], b[3]))
                    out[k] = out.get(k, 0.0) + e * a[4] * b[4]
    return out
