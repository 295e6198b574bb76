"""Generator consistency: the like-particle matrices assembled from the
per-parameter basis (one pass over the determinants for all parameters) equal
the matrices built directly for the same interaction, segment by segment."""
import numpy as np
import pytest

from nushellx_amd.synth.interaction import random_interaction
from nushellx_amd.synth.mspace import PF, SD, PNSpace
from nushellx_amd.synth.side import SideBasis


@pytest.mark.parametrize("orbs,kind,n", [(SD, "p", 2), (SD, "n", 3), (PF, "p", 2)])
def test_parameter_basis_reassembles_direct_segments(orbs, kind, n):
    inter = random_interaction(orbs, seed=11)
    side = SideBasis(PNSpace(orbs), kind, n)
    basis = side.hamiltonian_basis(inter)
    theta = inter.parameter_vector()
    for J2, (pairs, mats) in basis.items():
        direct = {(p, q): m for p, q, m in side.hamiltonian_segments(J2, inter)}
        for (p, q), M in zip(pairs, mats):
            got = np.tensordot(theta, M, axes=(0, 0))
            assert np.abs(got - direct[(p, q)]).max() < 1e-11, (J2, p, q)
        assert set(direct) <= set(pairs)
