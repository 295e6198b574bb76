"""Write a synthetic pn model into NuShellX's operator-file formats.

Files and record layouts follow the programs that produce them in the
reference pipeline:

* ``<base><J>.nus``  text input of NuTRL (``src/NuShellX/NuShellx.f90:41-70``,
  parsed by ``src/NuCommon/NuCommon.f90:2546-2736``);
* ``<base>a.tri/.trd``, ``<base>b.tri/.trd`` one-body densities per side
  (``src/NuOpd/NuOpd.f90:484-497,678-685``);
* ``<side><code>x.mat/.sgt`` like-particle H segments per side and J
  (``src/NuOrth/NuOrth.f90:220-223``; names ``src/NuCommon/NuModules.F90:553-597``);
* ``<base>x.oph`` pn multipole interaction (``src/NuOp/NuOp.f90:410-418``);
* ``<base>x.spe`` SPE / renorm / levels (``src/NuShellX/NuShellx.f90:171-187``).

Side roles follow NuShellX: the side with more valence nucleons (neutrons on
a tie) is "a" (``src/NuShellX/NuShellx.f90:569-580``).
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .fortio import RecordWriter
from .interaction import Interaction, multipole_pn
from .mspace import PNSpace
from .side import SideBasis

MAX_CODE = 72
AN_CODE = ("0123456789_$cdefghijklmnopqrstuvwxyz@ABCDEFGHIJKLMNMPQRSTUVWXYZ?a%^&=b-+#")
PTDIM = 32


def an_code(i: int) -> str:
    return AN_CODE[i]


def file_code(i: int, j: int, k: int) -> str:
    """3-character file code (``src/NuCommon/NuModules.F90:22-35``)."""
    f = lambda x: x + MAX_CODE if x < 0 else x
    return AN_CODE[f(i)] + AN_CODE[f(j)] + AN_CODE[f(k)]


def _fdiv(a: int, b: int) -> int:
    """Fortran integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def side_mat_name(side_nucleus: str, J2: int, T2: int, nmp: int, xtra: str) -> str:
    return side_nucleus + file_code(J2 // 2, T2 // 2, _fdiv(nmp, 2) + MAX_CODE // 2) + xtra


@dataclass
class PNModel:
    space: PNSpace
    nprot: int
    nneut: int
    inter: Interaction
    renorm: float = 1.0

    def roles(self) -> Tuple[str, str]:
        """(side a, side b) as 'p'/'n'."""
        return ("n", "p") if self.nneut >= self.nprot else ("p", "n")


def _partition_record(side: SideBasis) -> np.ndarray:
    arr = np.zeros((PTDIM, side.no_spart), dtype=np.int16)
    for p, occ in enumerate(side.partitions_global):
        arr[:len(occ), p] = occ
    return arr


def write_side_densities(side: SideBasis, path_base: str, xtag: float, max_lam: int,
                         dens: Optional[Dict] = None, real_kind: int = 4) -> None:
    """Write ``<path_base>.tri`` and ``<path_base>.trd`` (values REAL*4, or REAL*8
    as an rkw=8 build of the reference would write them)."""
    vdt = np.float32 if real_kind == 4 else np.float64
    if dens is None:
        dens = side.densities(max_lam)
    nsp = side.no_spart
    words = (nsp - 1) // 32 + 1
    minJ, maxJ = side.min_J2, side.max_J2
    with RecordWriter(path_base + ".tri") as w:
        w.write(nsp, minJ, maxJ, side.side == "p", side.side == "n", np.float32(xtag),
                0, b"a+a-", 2)
        w.write(side.nJT.astype(np.int32))
        w.write(_partition_record(side))
    # keys sorted as 8-byte strings (jl, jr, shl, shr, lam, 0, 0, 0)
    recs = []
    for (J2l, J2r, shl, shr, lam), D in dens.items():
        key = bytes([J2l, J2r, shl, shr, lam, 0, 0, 0])
        li = side.shells.index(shl)
        ri = side.shells.index(shr)
        jl_i, jr_i = side.jidx(J2l), side.jidx(J2r)
        offl = np.concatenate([[0], np.cumsum(side.nJT[:, jl_i])])
        offr = np.concatenate([[0], np.cumsum(side.nJT[:, jr_i])])
        maskcols = np.zeros((words, nsp), dtype=np.uint32)
        vals = []
        for pr in range(nsp):
            nr = side.nJT[pr, jr_i]
            if nr == 0:
                continue
            occ = list(side.partitions[pr])
            if occ[ri] == 0:
                continue
            occ[ri] -= 1
            occ[li] += 1
            if occ[li] > side.j2[li] + 1:
                continue
            pl = side.part_index.get(tuple(occ))
            if pl is None:
                continue
            nl = side.nJT[pl, jl_i]
            maskcols[pl // 32, pr] |= np.uint32(1 << (pl % 32))
            if nl == 0:
                continue
            vals.append(D[offl[pl]:offl[pl] + nl, offr[pr]:offr[pr] + nr].reshape(-1, order="F"))
        v = np.concatenate(vals).astype(vdt) if vals else np.zeros(0, vdt)
        if len(v) == 0:
            continue
        recs.append((key, v, maskcols))
    recs.sort(key=lambda r: r[0])
    total = sum(len(r[1]) for r in recs)
    with RecordWriter(path_base + ".trd") as w:
        w.write(len(recs), 2 * min(side.space.max_j2, side.space.max_j2), total, words, nsp, nsp, b"a+a-")
        if recs:
            w.write(b"".join(r[0] for r in recs))
            w.write(np.array([len(r[1]) for r in recs], dtype=np.int32))
            for key, v, mc in recs:
                w.write(v)
                w.write(mc.reshape(-1, order="F").view(np.int32))


def write_side_matrices(side: SideBasis, side_nucleus: str, outdir: str, inter: Interaction,
                        xtag: float, T2: int, nmp: int, xtra: str, real_kind: int = 4,
                        segments=None) -> None:
    """Write the ``.mat``/``.sgt`` pair for every J of a side (one file pair per J).
    ``segments`` optionally maps 2J to precomputed segments."""
    vdt = np.float32 if real_kind == 4 else np.float64
    for J2 in side.J2s:
        name = os.path.join(outdir, side_mat_name(side_nucleus, J2, T2, nmp, xtra))
        segs = segments[J2] if segments is not None else side.hamiltonian_segments(J2, inter)
        with RecordWriter(name + ".mat") as w:
            w.write(side.no_spart)
            for (np_, npr, m) in segs:
                occ = np.zeros(PTDIM, dtype=np.int16)
                occ[:side.space.no_shells] = side.partitions_global[np_ - 1]
                w.write(m.shape[1], m.shape[0], np_, npr, m.astype(vdt), 0, 0, occ)
        with RecordWriter(name + ".sgt") as w:
            w.write(side.no_spart)
            w.write(len(segs), side.dJ(J2), np.float32(xtag))


def write_nus(path: str, nucleus: str, space: PNSpace, nval: int, nmp: int, J2min: int,
              J2max: int, T2: int, tparity: int = 0, window: Tuple[int, int] = (0, 999)) -> None:
    """``window``: the N-hbar-omega limits (min_MSon, max_MSon); (0, 999) is none."""
    with open(path, "w") as f:
        f.write(f"{nucleus:>6s}\n")
        f.write("Synthetic pn model written by nushellx_amd\n")
        f.write("J-scheme operator files in NuShellX format\n")
        f.write("seeded interaction                      \n")
        f.write(f"{0:4d}{nval:4d}\n")
        f.write(f"{tparity:4d}\n")
        f.write(f"{space.no_shells:4d}\n")
        f.write("p" * space.k + "n" * space.k + "\n")
        for s in range(1, space.no_shells + 1):
            o = space.orbits[(s - 1) % space.k]
            f.write(f"{s:4d}{o.n:4d}{o.l:4d}{o.j2:4d}{1:4d}{o.parity:4d}\n")
        f.write(f"{J2min:4d}{J2max:4d}\n")
        f.write(f"{T2:4d}{T2:4d}{nmp:4d}\n")
        f.write(f"{window[0]:4d}{window[1]:4d}\n")
        f.write("".join(f"{0:4d}" for _ in range(space.no_shells)) + "\n")
        f.write("".join(f"{space.j2(s) + 1:4d}" for s in range(1, space.no_shells + 1)) + "\n")
        f.write("".join(f"{0:4d}" for _ in range(space.no_shells)) + "\n")
        f.write("".join(f"{99:4d}" for _ in range(space.no_shells)) + "\n")
        f.write(f"{0:4d}\n")
        f.write(f"{2:4d}{2:4d}\n")
        f.write("\n")


def write_oph(path: str, space: PNSpace, inter: Interaction) -> None:
    """Records (key, meV, pl, pr) with the pair parities pl = p1 p3 of the a-side
    pair and pr = p2 p4 of the b-side pair (``src/NuOp/NuOp.f90:378-379``)."""
    items, max_lmo = multipole_pn(space, inter)
    with RecordWriter(path) as w:
        w.write(len(items), max_lmo)
        for key, me in items:
            pl = space.parity(key[0]) * space.parity(key[1])
            pr = space.parity(key[2]) * space.parity(key[3])
            w.write(key, np.float64(me), pl, pr)


def write_spe(path: str, inter: Interaction, renorm: float, no_level: int,
              restart: Optional[int] = None) -> None:
    with open(path, "w") as f:
        spe = list(inter.spe) * 2
        f.write("".join(f"{e:10.4f}" for e in spe) + "\n")
        f.write(f"{renorm:14.10f}\n")
        f.write(f"{no_level:4d}{0:4d}{0:4d}\n")
        if restart is not None:
            f.write(f"{restart:4d}\n")


@dataclass
class ModelFiles:
    """Paths of one generated model (all J of one valence configuration)."""
    outdir: str
    base: str
    xtra: str
    J2s: Sequence[int]

    def nus(self, J2: int) -> str:
        return os.path.join(self.outdir, self.base + an_code(J2 // 2))

    @property
    def opfile_a(self) -> str:
        return os.path.join(self.outdir, self.base + "a")

    @property
    def opfile_b(self) -> str:
        return os.path.join(self.outdir, self.base + "b")

    @property
    def intfile(self) -> str:
        return os.path.join(self.outdir, self.base + self.xtra)

    def args(self, J2: int) -> Tuple[str, str, str, str]:
        """The four NuTRL command-line arguments (``src/NuShellX/NuShellx.f90:1060``)."""
        return (self.nus(J2), self.opfile_a, self.opfile_b, self.intfile)


class SideCache:
    """Interaction-independent side data (bases, densities) shared by many samples."""

    def __init__(self):
        self._sides: Dict[Tuple, SideBasis] = {}
        self._dens: Dict[Tuple, Dict] = {}
        self._basis: Dict[Tuple, Dict] = {}

    def side(self, space: PNSpace, kind: str, n: int) -> SideBasis:
        key = (space.orbits, kind, n)
        if key not in self._sides:
            self._sides[key] = SideBasis(space, kind, n)
        return self._sides[key]

    def segments(self, space: PNSpace, kind: str, n: int, inter: Interaction):
        """Like-particle H segments per 2J; linear in (SPE, TBME), so they are
        assembled from per-parameter basis matrices cached per side."""
        key = (space.orbits, kind, n)
        if key not in self._basis:
            self._basis[key] = self.side(space, kind, n).hamiltonian_basis(inter)
        basis = self._basis[key]
        theta = inter.parameter_vector()
        out = {}
        for J2, (pairs, mats) in basis.items():
            out[J2] = [(np_, npr, np.tensordot(theta, M, axes=(0, 0))) for (np_, npr), M in zip(pairs, mats)]
        return out

    def h_dense_basis(self, space: PNSpace, kind: str, n: int, inter: Interaction) -> Dict[int, np.ndarray]:
        """Per 2J the like-particle H(J) of each unit parameter as dense d x d
        matrices flattened column-major (P x d^2), in the partition order the
        .mat reader assembles them (block (npr, np) and its transpose)."""
        key = (space.orbits, kind, n)
        if key not in self._basis:
            self._basis[key] = self.side(space, kind, n).hamiltonian_basis(inter)
        side = self.side(space, kind, n)
        out = {}
        for J2, (pairs, mats) in self._basis[key].items():
            dims = [side.phi(J2)[p].shape[1] for p in range(side.no_spart)]
            off = np.concatenate([[0], np.cumsum(dims)]).astype(int)
            d = int(off[-1])
            P = mats[0].shape[0] if mats else len(inter.parameter_vector())
            D = np.zeros((P, d, d))
            for (np_, npr), M in zip(pairs, mats):
                p, q = np_ - 1, npr - 1
                D[:, off[q]:off[q] + dims[q], off[p]:off[p] + dims[p]] = M
                if p != q:
                    D[:, off[p]:off[p] + dims[p], off[q]:off[q] + dims[q]] = M.transpose(0, 2, 1)
            out[J2] = D.transpose(0, 2, 1).reshape(P, d * d)
        return out

    def dens(self, space: PNSpace, kind: str, n: int) -> Dict:
        key = (space.orbits, kind, n)
        if key not in self._dens:
            self._dens[key] = self.side(space, kind, n).densities(2 * space.max_j2)
        return self._dens[key]


def write_model(model: PNModel, outdir: str, base: str, J2s: Sequence[int], xtra: str = "x",
                no_level: int = 5, cache: Optional[SideCache] = None, seed_tag: int = 12345,
                densities: bool = True, real_kind: int = 4, tparity: int = 0,
                window: Tuple[int, int] = (0, 999)) -> ModelFiles:
    """Write every file NuTRL needs for the listed total 2J values.  ``tparity``
    (+1/-1, 0 for both) and ``window`` (N-hbar-omega limits) go into the .nus."""
    assert len(base) <= 5, "NuShellX nucleus names are at most 5 characters + J code"
    os.makedirs(outdir, exist_ok=True)
    cache = cache or SideCache()
    sp_ = model.space
    ra, rb = model.roles()
    nval = model.nprot + model.nneut
    nmp = model.nneut - model.nprot
    T2 = abs(nmp)
    rng = np.random.Generator(np.random.PCG64(seed_tag))
    for role, kind in (("a", ra), ("b", rb)):
        n = model.nprot if kind == "p" else model.nneut
        side = cache.side(sp_, kind, n)
        xtag = float(np.float32(rng.uniform(0.0, 1.0)))
        side_nuc = base + role
        if densities:
            write_side_densities(side, os.path.join(outdir, side_nuc), xtag, 2 * sp_.max_j2,
                                 dens=cache.dens(sp_, kind, n), real_kind=real_kind)
        else:
            # densities already on disk: keep their consistency tag
            from .fortio import read_records
            rec = next(read_records(os.path.join(outdir, side_nuc) + ".tri"))
            xtag = float(np.frombuffer(rec[20:24], dtype=np.float32)[0])
        side_nmp = -n if kind == "p" else n
        write_side_matrices(side, side_nuc, outdir, model.inter, xtag, T2, side_nmp, xtra,
                            real_kind=real_kind, segments=cache.segments(sp_, kind, n, model.inter))
    for J2 in J2s:
        write_nus(os.path.join(outdir, base + an_code(J2 // 2)) + ".nus", base + an_code(J2 // 2),
                  sp_, nval, nmp, J2, J2, T2, tparity=tparity, window=window)
    write_oph(os.path.join(outdir, base + xtra) + ".oph", sp_, model.inter)
    write_spe(os.path.join(outdir, base + xtra) + ".spe", model.inter, model.renorm, no_level)
    model.inter.write_int(os.path.join(outdir, base + xtra) + ".int")
    return ModelFiles(outdir, base, xtra, list(J2s))
