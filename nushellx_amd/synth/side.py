"""One side (protons or neutrons) of a pn J-scheme calculation, built clean-room.

For ``n`` valence nucleons of one kind this module constructs what NuShellX's
NuBasis/NuProj/NuMatrix/NuOrth/NuOpm/NuOpd chain leaves on disk for that side:

* the shell partitions in NuBasis order (Nijenhuis-Wilf compositions,
  ``src/NuBasis/NuBasis.f90:271-354``, filtered as ``:772-803``);
* nJT(p, J) = #(M=J) - #(M=J+1) Slater determinants of partition p;
* an orthonormal J basis per (partition, J): the kernel of J+ inside the M=J
  determinants (a deterministic construction; NuProj's is random-seeded, so
  only basis-invariant quantities are canonical);
* the like-particle Hamiltonian per J in that basis, SPE diagonal included
  (NuOrth, ``src/NuOrth/NuOrth.f90:206-218``), as partition segments;
* the reduced one-body densities in NuOpd's convention
  (``src/NuOpd/NuOpd.f90:422-449``; m-scheme phase ``src/NuOpm/NuOpm.f90:466``):
  rho = <Jl M=Jl| [a+_{sl} x a~_{sr}]^lam_{Jl-Jr} |Jr M=Jr> / <Jl Jl; Jr -Jr | lam Jl-Jr>
  with a~_{j,-m} = (-1)^{j-m} a_{j,m}.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from itertools import combinations
from typing import Dict, List, Tuple

import numpy as np
import scipy.linalg
import scipy.sparse as sp

from .angmom import cg
from .interaction import Interaction
from .mspace import PNSpace


def comp_next_all(n: int, k: int):
    """All compositions of n into k parts, in Nijenhuis-Wilf NEXCOM order."""
    a = [n] + [0] * (k - 1)
    t, h = n, 0
    out = [tuple(a)]
    while a[k - 1] != n:
        if t > 1:
            h = 0
        h += 1
        t = a[h - 1]
        a[h - 1] = 0
        a[0] = t - 1
        a[h] += 1
        out.append(tuple(a))
    return out


def _popcount(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    c = np.zeros(x.shape, dtype=np.int64)
    while np.any(x):
        c += (x & np.uint64(1)).astype(np.int64)
        x = x >> np.uint64(1)
    return c


def _below(bit: int) -> np.uint64:
    return np.uint64((1 << bit) - 1)


@dataclass
class Slice:
    """Determinants of one side with M = J (doubled), sorted by (partition, mask)."""
    masks: np.ndarray          # uint64
    part: np.ndarray           # partition index (0-based) per determinant
    start: np.ndarray          # per partition: first row of that partition in the slice
    count: np.ndarray          # per partition: number of determinants
    index: Dict[int, int] = field(default_factory=dict)


class SideBasis:
    """Slater determinants, partitions, J bases and operators of one side."""

    def __init__(self, space: PNSpace, side: str, nval: int):
        self.space = space
        self.side = side
        self.nval = nval
        self.shells = space.shells(side)          # global 1-based shell numbers
        self.j2 = [space.j2(s) for s in self.shells]
        # single-particle m-states, shell-major, m ascending
        ms_shell, ms_m2 = [], []
        for li, j2 in enumerate(self.j2):
            for m2 in range(-j2, j2 + 1, 2):
                ms_shell.append(li)
                ms_m2.append(m2)
        self.ms_shell = np.array(ms_shell)
        self.ms_m2 = np.array(ms_m2)
        self.nms = len(ms_shell)
        self.shell_first = [int(np.argmax(self.ms_shell == li)) for li in range(len(self.j2))]
        self._build_partitions()
        self._build_dets()

    # ---- partitions -------------------------------------------------------------
    def _build_partitions(self) -> None:
        sp_ = self.space
        tz = [-1 if s <= sp_.k else 1 for s in range(1, sp_.no_shells + 1)]
        nmp = -self.nval if self.side == "p" else self.nval
        parts = []
        for a in comp_next_all(self.nval, sp_.no_shells):
            if sum(x * t for x, t in zip(a, tz)) != nmp:
                continue
            if any(a[s - 1] > sp_.j2(s) + 1 for s in range(1, sp_.no_shells + 1)):
                continue
            parts.append(a)
        self.partitions_global = parts                       # occupation over all shells
        self.partitions = [tuple(p[s - 1] for s in self.shells) for p in parts]
        self.part_index = {p: i for i, p in enumerate(self.partitions)}
        self.no_spart = len(parts)
        self.part_parity = []
        for p in self.partitions:
            par = 1
            for li, occ in enumerate(p):
                par *= self.space.parity(self.shells[li]) ** occ
            self.part_parity.append(par)

    # ---- determinants --------------------------------------------------------------
    def _build_dets(self) -> None:
        masks = np.array([sum(1 << i for i in c) for c in combinations(range(self.nms), self.nval)],
                         dtype=np.uint64)
        m2 = np.zeros(len(masks), dtype=np.int64)
        occ = np.zeros((len(masks), len(self.j2)), dtype=np.int64)
        for i in range(self.nms):
            bit = ((masks >> np.uint64(i)) & np.uint64(1)).astype(np.int64)
            m2 += bit * self.ms_m2[i]
            occ[:, self.ms_shell[i]] += bit
        pidx = np.array([self.part_index.get(tuple(r), -1) for r in occ])
        self.det_masks, self.det_m2, self.det_part = masks, m2, pidx
        self.max_m2 = int(m2.max()) if len(m2) else 0
        self.min_J2 = self.nval % 2
        self.max_J2 = self.max_m2
        self.J2s = list(range(self.min_J2, self.max_J2 + 1, 2))
        self._slices: Dict[int, Slice] = {}
        self._phi: Dict[int, List[np.ndarray]] = {}
        self.nJT = np.zeros((self.no_spart, len(self.J2s)), dtype=np.int64)
        for jj, J2 in enumerate(self.J2s):
            s0, s1 = self.slice(J2), self.slice(J2 + 2)
            self.nJT[:, jj] = s0.count - s1.count

    def slice(self, M2: int) -> Slice:
        if M2 in self._slices:
            return self._slices[M2]
        sel = np.nonzero(self.det_m2 == M2)[0]
        order = np.lexsort((self.det_masks[sel], self.det_part[sel]))
        sel = sel[order]
        masks = self.det_masks[sel]
        part = self.det_part[sel]
        count = np.bincount(part, minlength=self.no_spart) if len(part) else np.zeros(self.no_spart, int)
        start = np.concatenate([[0], np.cumsum(count)[:-1]])
        s = Slice(masks, part, start, count, {int(m): i for i, m in enumerate(masks)})
        self._slices[M2] = s
        return s

    def jidx(self, J2: int) -> int:
        return (J2 - self.min_J2) // 2

    def dJ(self, J2: int) -> int:
        return int(self.nJT[:, self.jidx(J2)].sum())

    # ---- J basis ---------------------------------------------------------------------
    def _jplus(self, lo: Slice, hi: Slice) -> sp.csr_matrix:
        rows, cols, vals = [], [], []
        for i in range(self.nms):
            li = self.ms_shell[i]
            j2, m2 = self.j2[li], self.ms_m2[i]
            if m2 >= j2:
                continue
            b0, b1 = np.uint64(1 << i), np.uint64(1 << (i + 1))
            ok = ((lo.masks & b0) != 0) & ((lo.masks & b1) == 0)
            src = np.nonzero(ok)[0]
            if len(src) == 0:
                continue
            new = lo.masks[src] ^ b0 ^ b1
            dst = np.array([hi.index[int(x)] for x in new])
            coef = np.sqrt((j2 * (j2 + 2) - m2 * (m2 + 2)) / 4.0)
            rows.append(dst)
            cols.append(src)
            vals.append(np.full(len(src), coef))
        if not rows:
            return sp.csr_matrix((len(hi.masks), len(lo.masks)))
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(len(hi.masks), len(lo.masks)))

    def phi(self, J2: int) -> List[np.ndarray]:
        """Per partition: (#dets(p, M=J) x nJT(p, J)) orthonormal J=M states."""
        if J2 in self._phi:
            return self._phi[J2]
        lo, hi = self.slice(J2), self.slice(J2 + 2)
        jp = self._jplus(lo, hi).toarray()
        out = []
        for p in range(self.no_spart):
            a0, n0 = lo.start[p], lo.count[p]
            b0, n1 = hi.start[p], hi.count[p]
            want = int(self.nJT[p, self.jidx(J2)]) if J2 <= self.max_J2 else 0
            if n0 == 0 or want == 0:
                out.append(np.zeros((n0, 0)))
                continue
            block = jp[b0:b0 + n1, a0:a0 + n0]
            if n1 == 0:
                ns = np.eye(n0)
            else:
                ns = scipy.linalg.null_space(block, rcond=1e-10)
            # canonicalise: orthonormalise the projections of the unit vectors
            proj = ns @ ns.T
            basis = []
            for c in range(n0):
                v = proj[:, c].copy()
                for b in basis:
                    v -= b * (b @ v)
                for b in basis:
                    v -= b * (b @ v)
                nv = np.linalg.norm(v)
                if nv > 1e-6:
                    basis.append(v / nv)
                if len(basis) == want:
                    break
            if len(basis) != want:
                raise RuntimeError(f"J basis size {len(basis)} != nJT {want} (p={p}, 2J={J2})")
            B = np.array(basis).T
            # fixed sign: first significant component positive
            for c in range(B.shape[1]):
                k = int(np.argmax(np.abs(B[:, c]) > 1e-8))
                if B[k, c] < 0:
                    B[:, c] = -B[:, c]
            out.append(B)
        self._phi[J2] = out
        return out

    def phi_blockdiag(self, J2: int) -> sp.csr_matrix:
        return sp.block_diag([sp.csr_matrix(b) if b.size else sp.csr_matrix(b.shape)
                              for b in self.phi(J2)], format="csr")

    # ---- one-body operator a+_{il} a_{ir} between slices --------------------------------
    def _onebody(self, il: int, ir: int, L: Slice, R: Slice):
        bl, br = np.uint64(1 << il), np.uint64(1 << ir)
        has = (R.masks & br) != 0
        src = np.nonzero(has)[0]
        if len(src) == 0:
            return None
        m1 = R.masks[src] ^ br
        s1 = _popcount(R.masks[src] & _below(ir))
        free = (m1 & bl) == 0
        src, m1, s1 = src[free], m1[free], s1[free]
        if len(src) == 0:
            return None
        s2 = _popcount(m1 & _below(il))
        m2 = m1 | bl
        keep = np.array([int(x) in L.index for x in m2], dtype=bool)
        if not keep.any():
            return None
        src, m2, sgn = src[keep], m2[keep], ((s1 + s2)[keep] % 2)
        dst = np.array([L.index[int(x)] for x in m2])
        return dst, src, np.where(sgn == 1, -1.0, 1.0)

    # ---- reduced one-body densities ----------------------------------------------------
    def densities(self, max_lam: int):
        """Return {(2Jl, 2Jr, shl, shr, 2lam): dense dJ(Jl) x dJ(Jr) matrix} (global shells)."""
        out = {}
        for J2l in self.J2s:
            if self.dJ(J2l) == 0:
                continue
            Lsl = self.slice(J2l)
            PL = self.phi_blockdiag(J2l)
            for J2r in self.J2s:
                if self.dJ(J2r) == 0:
                    continue
                mu = J2l - J2r
                Rsl = self.slice(J2r)
                PR = self.phi_blockdiag(J2r)
                for lsl, jl in enumerate(self.j2):
                    for lsr, jr in enumerate(self.j2):
                        lams = [lam for lam in range(abs(jl - jr), min(jl + jr, max_lam) + 1, 2)
                                if abs(J2l - J2r) <= lam <= J2l + J2r and abs(mu) <= lam]
                        if not lams:
                            continue
                        T = {lam: None for lam in lams}
                        for ir in range(self.shell_first[lsr], self.shell_first[lsr] + jr + 1):
                            mr2 = int(self.ms_m2[ir])
                            ml2 = mr2 + mu
                            if abs(ml2) > jl:
                                continue
                            il = self.shell_first[lsl] + (ml2 + jl) // 2
                            e = self._onebody(il, ir, Lsl, Rsl)
                            if e is None:
                                continue
                            dst, src, sg = e
                            tilde = -1.0 if ((jr - mr2) // 2) % 2 else 1.0
                            E = sp.csr_matrix((sg * tilde, (dst, src)), shape=(len(Lsl.masks), len(Rsl.masks)))
                            for lam in lams:
                                c = cg(jl, jr, lam, ml2, -mr2, mu)
                                if c != 0.0:
                                    T[lam] = c * E if T[lam] is None else T[lam] + c * E
                        for lam in lams:
                            if T[lam] is None:
                                continue
                            c1 = cg(J2l, J2r, lam, J2l, -J2r, mu)
                            if c1 == 0.0:
                                continue
                            D = (PL.T @ (T[lam] @ PR)).toarray() / c1
                            out[(J2l, J2r, self.shells[lsl], self.shells[lsr], lam)] = D
        return out

    # ---- like-particle Hamiltonian -----------------------------------------------------
    def hamiltonian_m(self, M2: int, inter: Interaction) -> np.ndarray:
        """m-scheme H (SPE + T=1 two-body) in the M slice, dense."""
        S = self.slice(M2)
        n = len(S.masks)
        H = np.zeros((n, n))
        if n == 0:
            return H
        # one-body SPE
        for i in range(self.nms):
            occ = ((S.masks >> np.uint64(i)) & np.uint64(1)).astype(float)
            H[np.arange(n), np.arange(n)] += occ * inter.spe[self.ms_shell[i]]
        # two-body: V_as(ab, cd) for ordered pairs a<b, c<d of m-states
        pairs = [(a, b) for a in range(self.nms) for b in range(a + 1, self.nms)]
        by_m: Dict[int, List[Tuple[int, int]]] = {}
        for (a, b) in pairs:
            by_m.setdefault(int(self.ms_m2[a] + self.ms_m2[b]), []).append((a, b))
        o = inter.orbits
        jmax2 = 2 * max(x.j2 for x in o)
        for Mp, plist in by_m.items():
            # V matrix among pairs with this pair-M
            npair = len(plist)
            V = np.zeros((npair, npair))
            for x, (a, b) in enumerate(plist):
                oa, ob = self.ms_shell[a], self.ms_shell[b]
                for y, (c, d) in enumerate(plist):
                    oc, od = self.ms_shell[c], self.ms_shell[d]
                    s = 0.0
                    for J2 in range(0, jmax2 + 1, 2):
                        c1 = cg(o[oa].j2, o[ob].j2, J2, int(self.ms_m2[a]), int(self.ms_m2[b]), Mp)
                        if c1 == 0.0:
                            continue
                        c2 = cg(o[oc].j2, o[od].j2, J2, int(self.ms_m2[c]), int(self.ms_m2[d]), Mp)
                        if c2 == 0.0:
                            continue
                        nab = np.sqrt(2.0) if oa == ob else 1.0
                        ncd = np.sqrt(2.0) if oc == od else 1.0
                        s += c1 * c2 * nab * ncd * inter.v(oa, ob, oc, od, J2, 2)
                    V[x, y] = s
            for y, (c, d) in enumerate(plist):
                bc, bd = np.uint64(1 << c), np.uint64(1 << d)
                has = ((S.masks & bc) != 0) & ((S.masks & bd) != 0)
                src = np.nonzero(has)[0]
                if len(src) == 0:
                    continue
                m0 = S.masks[src]
                s1 = _popcount(m0 & _below(c))          # a_c
                m1 = m0 ^ bc
                s2 = _popcount(m1 & _below(d))          # a_d
                m2 = m1 ^ bd
                for x, (a, b) in enumerate(plist):
                    if V[x, y] == 0.0:
                        continue
                    ba, bb = np.uint64(1 << a), np.uint64(1 << b)
                    free = ((m2 & ba) == 0) & ((m2 & bb) == 0)
                    if not free.any():
                        continue
                    mm = m2[free]
                    s3 = _popcount(mm & _below(b))      # a+_b
                    m3 = mm | bb
                    s4 = _popcount(m3 & _below(a))      # a+_a
                    m4 = m3 | ba
                    sg = np.where(((s1[free] + s2[free] + s3 + s4) % 2) == 1, -1.0, 1.0)
                    dst = np.array([S.index[int(v)] for v in m4])
                    np.add.at(H, (dst, src[free]), sg * V[x, y])
        return H

    def hamiltonian_m_stack(self, M2: int, inter: Interaction) -> np.ndarray:
        """d H / d theta_i in the M slice for every parameter of
        ``Interaction.parameter_keys`` (dense, n_params x n x n).  The same
        second-quantised construction as ``hamiltonian_m``, with each pair
        matrix element kept as a vector over the parameters, so one pass over
        the determinants serves every parameter."""
        keys = inter.parameter_keys()
        pos = {k: i for i, k in enumerate(keys)}
        S = self.slice(M2)
        n = len(S.masks)
        H = np.zeros((len(keys), n, n))
        if n == 0:
            return H
        for i in range(self.nms):
            occ = ((S.masks >> np.uint64(i)) & np.uint64(1)).astype(float)
            H[pos[("spe", int(self.ms_shell[i]))], np.arange(n), np.arange(n)] += occ
        pairs = [(a, b) for a in range(self.nms) for b in range(a + 1, self.nms)]
        by_m: Dict[int, List[Tuple[int, int]]] = {}
        for (a, b) in pairs:
            by_m.setdefault(int(self.ms_m2[a] + self.ms_m2[b]), []).append((a, b))
        o = inter.orbits
        jmax2 = 2 * max(x.j2 for x in o)
        for Mp, plist in by_m.items():
            npair = len(plist)
            V = np.zeros((npair, npair, len(keys)))
            for x, (a, b) in enumerate(plist):
                oa, ob = self.ms_shell[a], self.ms_shell[b]
                for y, (c, d) in enumerate(plist):
                    oc, od = self.ms_shell[c], self.ms_shell[d]
                    for J2 in range(0, jmax2 + 1, 2):
                        c1 = cg(o[oa].j2, o[ob].j2, J2, int(self.ms_m2[a]), int(self.ms_m2[b]), Mp)
                        if c1 == 0.0:
                            continue
                        c2 = cg(o[oc].j2, o[od].j2, J2, int(self.ms_m2[c]), int(self.ms_m2[d]), Mp)
                        if c2 == 0.0:
                            continue
                        key, ph = inter.v_key(int(oa), int(ob), int(oc), int(od), J2, 2)
                        if ("v",) + key not in pos:
                            continue
                        nab = np.sqrt(2.0) if oa == ob else 1.0
                        ncd = np.sqrt(2.0) if oc == od else 1.0
                        V[x, y, pos[("v",) + key]] += c1 * c2 * nab * ncd * ph
            for y, (c, d) in enumerate(plist):
                bc, bd = np.uint64(1 << c), np.uint64(1 << d)
                has = ((S.masks & bc) != 0) & ((S.masks & bd) != 0)
                src = np.nonzero(has)[0]
                if len(src) == 0:
                    continue
                m0 = S.masks[src]
                s1 = _popcount(m0 & _below(c))
                m1 = m0 ^ bc
                s2 = _popcount(m1 & _below(d))
                m2 = m1 ^ bd
                for x, (a, b) in enumerate(plist):
                    nz = np.nonzero(V[x, y])[0]
                    if len(nz) == 0:
                        continue
                    ba, bb = np.uint64(1 << a), np.uint64(1 << b)
                    free = ((m2 & ba) == 0) & ((m2 & bb) == 0)
                    if not free.any():
                        continue
                    mm = m2[free]
                    s3 = _popcount(mm & _below(b))
                    m3 = mm | bb
                    s4 = _popcount(m3 & _below(a))
                    m4 = m3 | ba
                    sg = np.where(((s1[free] + s2[free] + s3 + s4) % 2) == 1, -1.0, 1.0)
                    dst = np.array([S.index[int(v)] for v in m4])
                    # (dst, src) pairs are distinct within one (x, y): plain indexed add
                    H[nz[:, None], dst[None, :], src[free][None, :]] += V[x, y, nz][:, None] * sg[None, :]
        return H

    def hamiltonian_basis(self, inter: Interaction):
        """Per 2J: (segment partition pairs, stack of per-parameter segment matrices)
        so that H(theta) segments = sum_i theta_i M_i (parameters of
        ``Interaction.parameter_keys``; T=0 TBMEs do not act between like nucleons)."""
        keys = inter.parameter_keys()
        out = {}
        for J2 in self.J2s:
            S = self.slice(J2)
            ph = self.phi(J2)
            pairs = [(p, q) for p in range(self.no_spart) if ph[p].shape[1]
                     for q in range(p, self.no_spart)
                     if ph[q].shape[1] and self.part_parity[p] == self.part_parity[q]]
            Hs = self.hamiltonian_m_stack(J2, inter)
            mats = []
            for (p, q) in pairs:
                blk = Hs[:, S.start[q]:S.start[q] + S.count[q], S.start[p]:S.start[p] + S.count[p]]
                mats.append(ph[q].T[None, :, :] @ blk @ ph[p][None, :, :])
            keep = [n for n, (p, q) in enumerate(pairs) if p == q or np.abs(mats[n]).max() > 1e-13]
            out[J2] = ([(pairs[n][0] + 1, pairs[n][1] + 1) for n in keep], [mats[n] for n in keep])
        return out

    def hamiltonian_segments(self, J2: int, inter: Interaction, tol: float = 1e-13):
        """Partition segments of H(J): list of (np, npr, m[dimptr x dimpt]), np<=npr, 1-based."""
        Hm = self.hamiltonian_m(J2, inter)
        S = self.slice(J2)
        ph = self.phi(J2)
        segs = []
        for p in range(self.no_spart):
            if ph[p].shape[1] == 0:
                continue
            for q in range(p, self.no_spart):
                if ph[q].shape[1] == 0 or self.part_parity[p] != self.part_parity[q]:
                    continue
                blk = Hm[S.start[q]:S.start[q] + S.count[q], S.start[p]:S.start[p] + S.count[p]]
                m = ph[q].T @ blk @ ph[p]
                if p == q or np.max(np.abs(m)) > tol:
                    segs.append((p + 1, q + 1, m))
        return segs
