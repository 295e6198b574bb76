"""Model spaces and proton-neutron shell numbering.

A pn model space lists every orbit twice: proton shells first (``nucleon='p'``),
then the same orbits as neutron shells (``'n'``), the layout NuShellX writes
into the ``.nus`` file (``src/NuShellX/NuShellx.f90:41-70``) and that
``no_shellsp``/``min_shellsa`` rely on (``src/NuCommonX/NuCommonX.f90:70-145``).
Shell numbers are 1-based, as in the operator-file keys.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple


@dataclass(frozen=True)
class Orbit:
    n: int
    l: int
    j2: int

    @property
    def parity(self) -> int:
        return -1 if self.l % 2 else 1

    @property
    def label(self) -> str:
        return f"{self.n}{'spdfghi'[self.l]}{self.j2}/2"


# orbit orders as stated for the synthetic configurations
SD = (Orbit(0, 2, 5), Orbit(1, 0, 1), Orbit(0, 2, 3))
PF = (Orbit(0, 3, 7), Orbit(1, 1, 3), Orbit(0, 3, 5), Orbit(1, 1, 1))

MODEL_SPACES = {"sd": SD, "pf": PF}


@dataclass(frozen=True)
class PNSpace:
    """Proton shells 1..k, neutron shells k+1..2k (same orbits)."""
    orbits: Tuple[Orbit, ...]

    @property
    def k(self) -> int:
        return len(self.orbits)

    @property
    def no_shells(self) -> int:
        return 2 * self.k

    def shells(self, side: str) -> List[int]:
        """1-based global shell numbers of a side ('p' or 'n')."""
        base = 0 if side == "p" else self.k
        return [base + i + 1 for i in range(self.k)]

    def j2(self, shell: int) -> int:
        return self.orbits[(shell - 1) % self.k].j2

    def orbit_index(self, shell: int) -> int:
        """Isospin-formalism orbit (0-based) of a pn shell."""
        return (shell - 1) % self.k

    def parity(self, shell: int) -> int:
        return self.orbits[(shell - 1) % self.k].parity

    @property
    def max_j2(self) -> int:
        return max(o.j2 for o in self.orbits)
