"""ORACLE (test infrastructure only): readers for NuShellX operator files and a
dense assembly of the pn J-scheme Hamiltonian by the reference algebra.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module.  It restates, for small cases, what NuTRL's setup
and H*v compute:

* block basis ``bas``/``posnv``/``dimnv`` (``src/NuTRL/NuCoreX.f90:932-988``):
  blocks (2Ja, 2Jb) with |Ja-Jb| <= J <= Ja+Jb, sorted ascending as 2-byte
  strings; inside a block the a index runs fastest;
* ``racahs`` (``:1071-1127``): r = (-1)^{(Ja+Jb'-J-lam)/2} W(Ja Jb Ja' Jb'; J lam) sqrt(lam+1);
* the drdbh term tables (``:1428-1577``);
* W_L = H_a V_L + V_L H_b^T + sum_{Ja', a-density k} A_k sum_{R, b-density k'}
  opmeV*r * V_R B_k'^T  (``multiplyx/a/b``, ``:76-226``; ``src/NuCommonX/NuCommonX.f90:745-1104``).

Dense matrices are built block by block so the result is the dense H that
NuTRL's ``matrixs`` (``:1370-1401``) would dump, without its REAL*4 rounding
and without the 1e-7 product thresholds (an option restores them).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass
from typing import Dict, List, Tuple

import numpy as np

MAX_CODE = 72
AN = "0123456789_$cdefghijklmnopqrstuvwxyz@ABCDEFGHIJKLMNMPQRSTUVWXYZ?a%^&=b-+#"


def records(path: str):
    data = open(path, "rb").read()
    pos = 0
    while pos < len(data):
        n = struct.unpack_from("<i", data, pos)[0]
        yield data[pos + 4:pos + 4 + n]
        pos += n + 8


def _fdiv(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def fcode(i, j, k):
    f = lambda x: x + MAX_CODE if x < 0 else x
    return AN[f(i)] + AN[f(j)] + AN[f(k)]


# ---------------------------------------------------------------- 6j (binomial form)
def _binm(n, k):
    from math import comb
    return float(comb(n, k)) if 0 <= k <= n else 0.0


def _fact(n):
    from math import factorial
    return float(factorial(n))


def _angdelta(a, b, c):
    return np.sqrt(_fact((a + b - c) // 2) / _fact((a + b + c) // 2 + 1)
                   * _fact((a - b + c) // 2) * _fact((-a + b + c) // 2))


def sixj(a, b, c, d, e, f):
    """{a b c; d e f}, doubled args; binomial-sum form as ``src/NuCommon/NuCommon.f90:3268-3309``."""
    for (x, y, z) in ((a, b, c), (c, d, e), (a, e, f), (b, d, f)):
        if (x + y - z) % 2 or abs(x - y) > z or x + y < z:
            return 0.0
    out = _angdelta(a, e, f) / _angdelta(a, b, c) * _angdelta(b, d, f) * _angdelta(c, d, e)
    nlo = max((a + b + c) // 2, (c + d + e) // 2, (b + d + f) // 2, (a + e + f) // 2)
    nhi = min((a + b + d + e) // 2, (b + c + e + f) // 2, (a + c + d + f) // 2)
    s = 0.0
    for n in range(nlo, nhi + 1):
        s += ((-1) ** n * _binm(n + 1, n - (a + b + c) // 2) * _binm((a + b - c) // 2, n - (c + d + e) // 2)
              * _binm((a - b + c) // 2, n - (b + d + f) // 2) * _binm((b - a + c) // 2, n - (a + e + f) // 2))
    return s * out


def racah(j1, j2, j3, j4, J, K):
    ph = -1.0 if ((j1 + j2 + j3 + j4) // 2) % 2 else 1.0
    return ph * sixj(j1, j2, J, j4, j3, K)


# ---------------------------------------------------------------- file model
@dataclass
class Nus:
    nucleus: str
    no_cI: int
    no_cN: int
    tparity: int
    no_shells: int
    nucleon: str
    sj2: List[int]
    sp: List[int]
    min_cJ2: int
    max_cJ2: int
    min_cT2: int
    max_cT2: int
    nmp: int
    min_MSon: int
    max_MSon: int
    del_cJ2: int
    del_cT2: int


def read_nus(path: str) -> Nus:
    lines = [l.rstrip("\n") for l in open(path)]
    lines = [l for l in lines if not l.startswith("!")]
    it = iter(lines)
    nuc = next(it).strip()
    for _ in range(3):
        next(it)
    no_cI, no_cN = (int(x) for x in next(it).split()[:2])
    tpar = int(next(it).split()[0])
    nsh = int(next(it).split()[0])
    nucleon = next(it).strip()
    sj2, spar = [], []
    for _ in range(nsh):
        p = [int(x) for x in next(it).split()]
        sj2.append(p[3]); spar.append(p[5])
    mnJ, mxJ = (int(x) for x in next(it).split()[:2])
    mnT, mxT, nmp = (int(x) for x in next(it).split()[:3])
    mnMS, mxMS = (int(x) for x in next(it).split()[:2])
    for _ in range(4):
        next(it)
    next(it)  # output_control
    dJ, dT = 2, 2
    try:
        p = next(it).split()
        if len(p) >= 2:
            dJ, dT = int(p[0]), int(p[1])
    except StopIteration:
        pass
    if dJ < 2 or dJ % 2:
        dJ = 2
    if dT < 2 or dT % 2:
        dT = 2
    return Nus(nuc, no_cI, no_cN, tpar, nsh, nucleon, sj2, spar, mnJ, mxJ, mnT, mxT, nmp, mnMS, mxMS, dJ, dT)


@dataclass
class SideOps:
    no_spart: int
    minJ2: int
    maxJ2: int
    proton: bool
    neutron: bool
    xtag: float
    delJ2: int
    nJT: np.ndarray            # (no_spart, nJ)
    spart: np.ndarray          # (no_spart, 32) int16
    dens: Dict[bytes, Tuple[np.ndarray, np.ndarray]]   # key -> (values f32, mask words (nsp, words))
    max_lm: int

    def nJ(self, J2):
        return self.nJT[:, J2 // 2]

    def d(self, J2):
        return int(self.nJT[:, J2 // 2].sum())


def read_side(opfile: str) -> SideOps:
    r = list(records(opfile + ".tri"))
    nsp, mnJ, mxJ, pr, ne = struct.unpack_from("<5i", r[0], 0)
    x = struct.unpack_from("<f", r[0], 20)[0]
    dJ = struct.unpack_from("<i", r[0], 32)[0]
    nJ = mxJ // 2 - mnJ // 2 + 1
    nJT = np.zeros((nsp, mxJ // 2 + 1), dtype=np.int64)
    nJT[:, mnJ // 2:] = np.frombuffer(r[1], dtype=np.int32)[:nsp * nJ].reshape(nsp, nJ, order="F")
    spart = np.frombuffer(r[2], dtype=np.int16).reshape(nsp, -1)
    t = list(records(opfile + ".trd"))
    trdd, max_lm, tot = struct.unpack_from("<3i", t[0], 0)
    words = (nsp - 1) // 32 + 1
    dens = {}
    if trdd:
        keys = [t[1][8 * i:8 * i + 8] for i in range(trdd)]
        cnt = np.frombuffer(t[2], dtype=np.int32)
        pos = 3
        for k, c in zip(keys, cnt):
            if c == 0:
                dens[k] = (np.zeros(0, np.float32), np.zeros((nsp, words), np.uint32))
                continue
            v = np.frombuffer(t[pos], dtype=np.float32)[:c]
            m = np.frombuffer(t[pos + 1], dtype=np.uint32)[:nsp * words].reshape(nsp, words)
            dens[k] = (v, m)
            pos += 2
    return SideOps(nsp, mnJ, mxJ, pr != 0, ne != 0, x, dJ, nJT, spart, dens, max_lm)


def read_mats(side_nucleus: str, side: SideOps, minT2: int, maxT2: int, nmp: int, xtra: str):
    """Dense H per J (float64 from REAL*4 segments), from ``.mat``/``.sgt``."""
    sp_par = None
    out = {}
    for J2 in range(side.minJ2, side.maxJ2 + 1, 2):
        for T2 in range(minT2, maxT2 + 1, 2):
            name = side_nucleus + fcode(J2 // 2, T2 // 2, _fdiv(nmp, 2) + MAX_CODE // 2) + xtra
            if not os.path.exists(name + ".sgt"):
                continue
            sg = list(records(name + ".sgt"))
            if len(sg) < 2:
                continue
            nseg, mdim = struct.unpack_from("<2i", sg[1], 0)
            xx = struct.unpack_from("<f", sg[1], 8)[0]
            if xx != side.xtag:
                raise ValueError("mat file and proj file not consistent")
            d = side.d(J2)
            off = np.concatenate([[0], np.cumsum(side.nJ(J2))])
            H = np.zeros((d, d))
            mr = list(records(name + ".mat"))[1:1 + nseg]
            for rec in mr:
                dimpt, dimptr, npc, npr = struct.unpack_from("<4i", rec, 0)
                m = np.frombuffer(rec, dtype=np.float32, count=dimpt * dimptr, offset=16)
                m = m.astype(np.float64).reshape(dimptr, dimpt, order="F")
                r0, c0 = off[npr - 1], off[npc - 1]
                H[r0:r0 + dimptr, c0:c0 + dimpt] = m
                if npr != npc:
                    H[c0:c0 + dimpt, r0:r0 + dimptr] = m.T
            out[J2] = H
    return out


def partition_parity(spart_row, sp):
    par = 1
    for i, occ in enumerate(spart_row[:len(sp)]):
        if occ > 0:
            par *= sp[i] ** int(occ)
    return par


def dense_density(side: SideOps, key: bytes) -> np.ndarray:
    """Dense d(Jl) x d(Jr) matrix of one density, walking bits column-major as next_ij."""
    J2l, J2r = key[0], key[1]
    vals, mask = side.dens[key]
    nl, nr = side.nJ(J2l), side.nJ(J2r)
    offl = np.concatenate([[0], np.cumsum(nl)])
    offr = np.concatenate([[0], np.cumsum(nr)])
    D = np.zeros((side.d(J2l), side.d(J2r)))
    pt = 0
    for j in range(side.no_spart):            # column partition (right)
        for i in range(side.no_spart):        # row partition (left)
            if (int(mask[j, i // 32]) >> (i % 32)) & 1:
                blk = vals[pt:pt + nl[i] * nr[j]].astype(np.float64).reshape(nl[i], nr[j], order="F")
                D[offl[i]:offl[i] + nl[i], offr[j]:offr[j] + nr[j]] = blk
                pt += nl[i] * nr[j]
    return D


@dataclass
class PNJ:
    J2: int
    blocks: List[Tuple[int, int]]      # (2Ja, 2Jb) in bas order
    posnv: List[int]
    dimnv: List[int]
    N: int
    Ha: Dict[int, np.ndarray]
    Hb: Dict[int, np.ndarray]
    a: SideOps
    b: SideOps
    terms: list                         # (L, Ja', A dense, [(R, coef, B dense)])


def load(nus_prefix: str, opfile_a: str, opfile_b: str, intfile: str, thresholds: bool = False) -> PNJ:
    """Read NuTRL's inputs (``src/NuTRL/NuCoreX.f90:1634-1670``) and build the tables."""
    nus = read_nus(nus_prefix + ".nus")
    a = read_side(opfile_a)
    b = read_side(opfile_b)
    typea = "p" if a.proton else "n"
    typeb = "p" if b.proton else "n"
    J2 = nus.min_cJ2
    np_ = nus.nucleon.count("p")
    shells_p = list(range(1, np_ + 1))
    shells_n = list(range(np_ + 1, nus.no_shells + 1))
    sh_a = shells_p if typea == "p" else shells_n
    sh_b = shells_p if typeb == "p" else shells_n
    if typea == "p":
        no_cB, no_cA = (nus.no_cN + nus.nmp) // 2, (nus.no_cN - nus.nmp) // 2
    else:
        no_cB, no_cA = (nus.no_cN - nus.nmp) // 2, (nus.no_cN + nus.nmp) // 2
    # blocks
    blocks = []
    for ajl in range(a.minJ2, a.maxJ2 + 1, a.delJ2):
        bjln = max(abs(ajl - J2), b.minJ2)
        bjlx = min(ajl + J2, b.maxJ2)
        bjln = bjln + (bjln - b.minJ2) % b.delJ2
        for bjl in range(bjln, bjlx + 1, b.delJ2):
            blocks.append((ajl, bjl))
    blocks.sort()
    posnv, dimnv, N = [], [], 0
    for (ja, jb) in blocks:
        posnv.append(N)
        dimnv.append(a.d(ja) * b.d(jb))
        N += dimnv[-1]
    # spe / renorm
    sl = open(intfile + ".spe").read().splitlines()
    renorm = float(sl[1].split()[0])
    # oph
    op = list(records(intfile + ".oph"))
    opd, max_lmo = struct.unpack_from("<2i", op[0], 0)
    opn = {}
    for rec in op[1:1 + opd]:
        k = rec[:8]
        v = struct.unpack_from("<d", rec, 8)[0]
        if abs(v) < 1e-10:
            v = 0.0
        opn[k] = renorm * v
    # like-particle matrices
    base_a = os.path.join(os.path.dirname(opfile_a), os.path.basename(opfile_a))
    xtra = os.path.basename(intfile)[-1]
    nmp_a = no_cA if typea == "n" else -no_cA
    nmp_b = no_cB if typeb == "n" else -no_cB
    Ha = read_mats(opfile_a, a, nus.min_cT2, nus.max_cT2, nmp_a, xtra)
    Hb = read_mats(opfile_b, b, nus.min_cT2, nus.max_cT2, nmp_b, xtra)
    # lambda limits (``:1320-1326``)
    max_sj2u = max(nus.sj2)
    max_lambda = nus.del_cJ2
    if max_lambda <= 2 or max_lambda % 2:
        max_lambda = 2 * max_sj2u
    del_lm = nus.del_cT2
    if del_lm < 2 or del_lm % 2:
        del_lm = 2
    max_lm = min(a.max_lm, max_lmo, 2 * max_sj2u, max_lambda)
    bidx = {blk: i for i, blk in enumerate(blocks)}
    terms = []
    for L, (ajl, bjl) in enumerate(blocks):
        for ajr in range(a.minJ2, a.maxJ2 + 1, a.delJ2):
            for ashl in sh_a:
                for ashr in sh_a:
                    j1, j2 = nus.sj2[ashl - 1], nus.sj2[ashr - 1]
                    lmn = abs(j1 - j2)
                    lmn += lmn % del_lm
                    lmx = min(j1 + j2, max_lm)
                    for lm in range(lmn, lmx + 1, del_lm):
                        if not (abs(ajl - lm) <= ajr <= ajl + lm):
                            continue
                        ta = bytes([ajl, ajr, ashl, ashr, lm, 0, 0, 0])
                        if ta not in a.dens:
                            continue
                        rlist = []
                        bjrn = max(abs(bjl - lm), abs(ajr - J2))
                        bjrx = min(bjl + lm, ajr + J2, b.maxJ2)
                        bjrn = bjrn + (bjrn - b.minJ2) % b.delJ2
                        for bjr in range(bjrn, bjrx + 1, b.delJ2):
                            rc = racah(ajl, bjl, ajr, bjr, J2, lm)
                            if abs(rc) == 0.0:
                                continue
                            ph = -1.0 if ((ajl + bjr - J2 - lm) // 2) % 2 else 1.0
                            r = ph * rc * np.sqrt(lm + 1.0)
                            R = bidx.get((ajr, bjr))
                            if R is None:
                                raise ValueError("find error bh tables")
                            for bshl in sh_b:
                                for bshr in sh_b:
                                    k1, k2 = nus.sj2[bshl - 1], nus.sj2[bshr - 1]
                                    if lm < abs(k1 - k2) or lm > min(k1 + k2, max_lm):
                                        continue
                                    if not (abs(bjl - lm) <= bjr <= bjl + lm):
                                        continue
                                    tb = bytes([bjl, bjr, bshl, bshr, lm, 0, 0, 0])
                                    if tb not in b.dens:
                                        continue
                                    opk = bytes([ashl, ashr, bshl, bshr, lm, 0, 0, 0])
                                    me = opn.get(opk, 0.0)
                                    if me == 0.0:
                                        continue
                                    rlist.append((R, me * r, tb))
                        terms.append((L, ajr, ta, rlist))
    pnj = PNJ(J2, blocks, posnv, dimnv, N, Ha, Hb, a, b, terms)
    pnj.thresholds = thresholds
    return pnj


def dense_h(p: PNJ) -> np.ndarray:
    """Dense H assembled block by block (brute-force oracle, small N only)."""
    H = np.zeros((p.N, p.N))
    for L, (ja, jb) in enumerate(p.blocks):
        da, db = p.a.d(ja), p.b.d(jb)
        s = slice(p.posnv[L], p.posnv[L] + da * db)
        # vec(A V B^T) = (B kron A) vec(V) for column-major vec
        if ja in p.Ha:
            H[s, s] += np.kron(np.eye(db), p.Ha[ja])
        if jb in p.Hb:
            H[s, s] += np.kron(p.Hb[jb], np.eye(da))
    cacheA, cacheB = {}, {}
    for (L, ajr, ta, rlist) in p.terms:
        if ta not in cacheA:
            cacheA[ta] = dense_density(p.a, ta)
        A = cacheA[ta]
        jb = p.blocks[L][1]
        for (R, c, tb) in rlist:
            if tb not in cacheB:
                cacheB[tb] = dense_density(p.b, tb)
            B = cacheB[tb]
            cB = c * B
            if getattr(p, "thresholds", False):
                cB = np.where(np.abs(cB) >= 1e-7, cB, 0.0)
                A = np.where(np.abs(A) >= 1e-7, A, 0.0)
            sL = slice(p.posnv[L], p.posnv[L] + p.dimnv[L])
            sR = slice(p.posnv[R], p.posnv[R] + p.dimnv[R])
            H[sL, sR] += np.kron(cB, A)
    return H
