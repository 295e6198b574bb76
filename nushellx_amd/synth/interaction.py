"""Seeded isospin-conserving two-body interactions and their pn multipole form.

* ``random_interaction`` draws the J-T-coupled TBMEs <ab;JT|V|cd;JT>
  (a<=b, c<=d, (ab)<=(cd), Pauli-allowed) i.i.d. N(0, sigma) and the SPEs
  uniform on [lo, hi], from ``numpy.random.Generator(PCG64(seed))``.
* ``Interaction.v(a, b, c, d, J, T)`` returns the element for any orbit order,
  with the exchange phase (-1)^{(ja+jb-J-T)} of ``findV``
  (``src/NuOp/NuOp.f90:448-532``) and bra-ket symmetry.
* ``multipole_pn`` is the particle-hole recoupling NuOp writes to ``.oph``
  (``src/NuOp/NuOp.f90:352-363``):
  V^lam(a a' ; b b') = sum_{J,T} V_JT(j1 j2; j3 j4) (J+1) sqrt(lam+1)
                       (-1)^{(j1+j4-J-lam)/2} W(j1 j2 j3 j4; J lam) d12 d34 / 2
  with j1=a (created, side a), j3=a' (annihilated, side a), j2=b, j4=b'.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from .angmom import racah
from .mspace import Orbit, PNSpace

Key = Tuple[int, int, int, int, int, int]  # (a, b, c, d, 2J, 2T), orbits 0-based


def _allowed(orbits, a, b, J2, T2) -> bool:
    ja, jb = orbits[a].j2, orbits[b].j2
    if not (abs(ja - jb) <= J2 <= ja + jb):
        return False
    if a == b and ((J2 // 2 + T2 // 2) % 2 == 0):
        return False  # J+T must be odd for two nucleons in one orbit
    return True


def tbme_keys(orbits: Tuple[Orbit, ...]) -> List[Key]:
    """Canonical list of independent JT matrix elements (parity-conserving)."""
    n = len(orbits)
    pairs = [(a, b) for a in range(n) for b in range(a, n)]
    keys: List[Key] = []
    for T2 in (0, 2):
        for i, (a, b) in enumerate(pairs):
            for (c, d) in pairs[i:]:
                if orbits[a].parity * orbits[b].parity != orbits[c].parity * orbits[d].parity:
                    continue
                for J2 in range(0, 2 * max(o.j2 for o in orbits) + 1, 2):
                    if _allowed(orbits, a, b, J2, T2) and _allowed(orbits, c, d, J2, T2):
                        keys.append((a, b, c, d, J2, T2))
    return keys


@dataclass
class Interaction:
    orbits: Tuple[Orbit, ...]
    spe: np.ndarray                      # per orbit (MeV)
    tbme: Dict[Key, float] = field(default_factory=dict)

    def v_key(self, a: int, b: int, c: int, d: int, J2: int, T2: int):
        """(canonical key, phase) with <ab;JT|V|cd;JT> = phase * tbme[key]."""
        o = self.orbits
        ph = 1.0
        if a > b:
            if ((o[a].j2 + o[b].j2 - J2 - T2) // 2) % 2:
                ph = -ph
            a, b = b, a
        if c > d:
            if ((o[c].j2 + o[d].j2 - J2 - T2) // 2) % 2:
                ph = -ph
            c, d = d, c
        if (a, b) > (c, d):
            a, b, c, d = c, d, a, b
        return (a, b, c, d, J2, T2), ph

    def v(self, a: int, b: int, c: int, d: int, J2: int, T2: int) -> float:
        """<ab;JT|V|cd;JT> (normalised, antisymmetrised) for any orbit order."""
        o = self.orbits
        ph = 1.0
        if a > b:
            if ((o[a].j2 + o[b].j2 - J2 - T2) // 2) % 2:
                ph = -ph
            a, b = b, a
        if c > d:
            if ((o[c].j2 + o[d].j2 - J2 - T2) // 2) % 2:
                ph = -ph
            c, d = d, c
        if (a, b) > (c, d):
            a, b, c, d = c, d, a, b
        return ph * self.tbme.get((a, b, c, d, J2, T2), 0.0)

    # ---- linear parametrisation: H = sum_i theta_i H_i over (SPEs, TBMEs) ----------
    def parameter_keys(self) -> List[tuple]:
        return [("spe", o) for o in range(len(self.orbits))] + [("v",) + k for k in sorted(self.tbme)]

    def parameter_vector(self) -> np.ndarray:
        return np.concatenate([np.asarray(self.spe, dtype=float),
                               np.array([self.tbme[k] for k in sorted(self.tbme)], dtype=float)])

    def unit(self, i: int) -> "Interaction":
        """The interaction with parameter i equal to 1 and every other parameter 0."""
        keys = sorted(self.tbme)
        spe = np.zeros(len(self.orbits))
        tb = {k: 0.0 for k in keys}
        if i < len(spe):
            spe[i] = 1.0
        else:
            tb[keys[i - len(spe)]] = 1.0
        return Interaction(self.orbits, spe, tb)

    def perturbed(self, seed: int, rel_sigma: float = 0.1) -> "Interaction":
        """V_s = V_0 + eps, eps_i ~ N(0, (rel_sigma |V_0,i|)^2) independently."""
        rng = np.random.Generator(np.random.PCG64(seed))
        keys = sorted(self.tbme)
        vals = np.array([self.tbme[k] for k in keys])
        eps = rng.normal(0.0, 1.0, size=len(keys)) * rel_sigma * np.abs(vals)
        return Interaction(self.orbits, self.spe.copy(), dict(zip(keys, vals + eps)))

    def write_int(self, path: str) -> None:
        """Text interaction file: SPE line, then 'a b c d 2J 2T V' (1-based orbits)."""
        with open(path, "w") as f:
            f.write(f"{len(self.tbme)} " + " ".join(f"{e:.12f}" for e in self.spe) + "\n")
            for (a, b, c, d, J2, T2) in sorted(self.tbme):
                f.write(f"{a+1} {b+1} {c+1} {d+1} {J2//2} {T2//2} {self.tbme[(a,b,c,d,J2,T2)]:.15f}\n")


def random_interaction(orbits: Tuple[Orbit, ...], seed: int, sigma: float = 1.5,
                       spe_range: Tuple[float, float] = (-4.0, 0.0)) -> Interaction:
    rng = np.random.Generator(np.random.PCG64(seed))
    keys = tbme_keys(orbits)
    vals = rng.normal(0.0, sigma, size=len(keys))
    spe = rng.uniform(spe_range[0], spe_range[1], size=len(orbits))
    return Interaction(tuple(orbits), spe, dict(zip(keys, vals)))


def read_int(path: str, orbits: Tuple[Orbit, ...]) -> Interaction:
    with open(path) as f:
        head = f.readline().split()
        spe = np.array([float(x) for x in head[1:]])
        tb: Dict[Key, float] = {}
        for line in f:
            p = line.split()
            if len(p) < 7:
                continue
            a, b, c, d, J, T = (int(x) for x in p[:6])
            tb[(a - 1, b - 1, c - 1, d - 1, 2 * J, 2 * T)] = float(p[6])
    return Interaction(tuple(orbits), spe, tb)


def multipole_pn(space: PNSpace, inter: Interaction):
    """All .oph records: list of (key bytes, meV) in ascending key order, and max_lmo.

    Both assignments of the proton and neutron sides to the solver's a/b roles
    are written, as NuOp does (``pnnp`` loop, ``src/NuOp/NuOp.f90:247-270``).
    """
    o = inter.orbits
    recs: Dict[bytes, float] = {}
    max_lmo = 0
    for a_side, b_side in (("p", "n"), ("n", "p")):
        for sh1 in space.shells(a_side):
            j1 = space.j2(sh1)
            o1 = space.orbit_index(sh1)
            for sh2 in space.shells(b_side):
                j2 = space.j2(sh2)
                o2 = space.orbit_index(sh2)
                d12 = np.sqrt(2.0) if o1 == o2 else 1.0
                for sh3 in space.shells(a_side):
                    j3 = space.j2(sh3)
                    o3 = space.orbit_index(sh3)
                    for sh4 in space.shells(b_side):
                        j4 = space.j2(sh4)
                        o4 = space.orbit_index(sh4)
                        d34 = np.sqrt(2.0) if o3 == o4 else 1.0
                        lmn = max(abs(j1 - j3), abs(j2 - j4))
                        lmx = min(j1 + j3, j2 + j4)
                        max_lmo = max(max_lmo, lmx)
                        for lam in range(lmx, lmn - 1, -2):
                            s = 0.0
                            for J2 in range(max(abs(j1 - j2), abs(j3 - j4)),
                                            min(j1 + j2, j3 + j4) + 1, 2):
                                for T2 in (0, 2):
                                    if (o1 == o2 or o3 == o4):
                                        # two nucleons in one orbit: J+T odd only
                                        if ((J2 // 2 + T2 // 2) % 2) == 0:
                                            continue
                                    V = inter.v(o1, o2, o3, o4, J2, T2)
                                    if V == 0.0:
                                        continue
                                    ph = -1.0 if ((j1 + j4 - J2 - lam) // 2) % 2 else 1.0
                                    s += (V * (J2 + 1) * np.sqrt(lam + 1.0) * ph
                                          * racah(j1, j2, j3, j4, J2, lam) * d12 * d34 / 2.0)
                            key = bytes([sh1, sh3, sh2, sh4, lam, 0, 0, 0])
                            if key not in recs:
                                recs[key] = s
    items = sorted(recs.items())
    return items, max_lmo


def multipole_pn_basis(space: PNSpace, inter: Interaction):
    """The .oph records as linear functions of the interaction parameters: keys
    (ascending, as ``multipole_pn`` writes them), max_lmo, and M (P x records)
    with ``multipole_pn(space, V)`` values = theta(V) @ M for every interaction V
    with the TBME keys of ``inter`` (parameters of ``Interaction.parameter_keys``;
    the SPE rows are zero)."""
    o = inter.orbits
    pkeys = sorted(inter.tbme)
    pidx = {k: len(inter.spe) + i for i, k in enumerate(pkeys)}
    P = len(inter.spe) + len(pkeys)
    recs: Dict[bytes, Dict[int, float]] = {}
    max_lmo = 0
    for a_side, b_side in (("p", "n"), ("n", "p")):
        for sh1 in space.shells(a_side):
            j1, o1 = space.j2(sh1), space.orbit_index(sh1)
            for sh2 in space.shells(b_side):
                j2, o2 = space.j2(sh2), space.orbit_index(sh2)
                d12 = np.sqrt(2.0) if o1 == o2 else 1.0
                for sh3 in space.shells(a_side):
                    j3, o3 = space.j2(sh3), space.orbit_index(sh3)
                    for sh4 in space.shells(b_side):
                        j4, o4 = space.j2(sh4), space.orbit_index(sh4)
                        d34 = np.sqrt(2.0) if o3 == o4 else 1.0
                        lmn = max(abs(j1 - j3), abs(j2 - j4))
                        lmx = min(j1 + j3, j2 + j4)
                        max_lmo = max(max_lmo, lmx)
                        for lam in range(lmx, lmn - 1, -2):
                            key = bytes([sh1, sh3, sh2, sh4, lam, 0, 0, 0])
                            if key in recs:
                                continue
                            col: Dict[int, float] = {}
                            for J2 in range(max(abs(j1 - j2), abs(j3 - j4)), min(j1 + j2, j3 + j4) + 1, 2):
                                for T2 in (0, 2):
                                    if (o1 == o2 or o3 == o4) and ((J2 // 2 + T2 // 2) % 2) == 0:
                                        continue
                                    k, phv = inter.v_key(o1, o2, o3, o4, J2, T2)
                                    if k not in pidx:
                                        continue
                                    ph = -1.0 if ((j1 + j4 - J2 - lam) // 2) % 2 else 1.0
                                    c = (phv * (J2 + 1) * np.sqrt(lam + 1.0) * ph
                                         * racah(j1, j2, j3, j4, J2, lam) * d12 * d34 / 2.0)
                                    col[pidx[k]] = col.get(pidx[k], 0.0) + c
                            recs[key] = col
    keys = sorted(recs)
    M = np.zeros((P, len(keys)))
    for r, k in enumerate(keys):
        for i, c in recs[k].items():
            M[i, r] += c
    return keys, max_lmo, M
