"""Linear-ensemble inputs on the CPU: the generator's linear forms of the .oph
multipoles and the like-particle H(J), checked against the values the files of
an interaction actually hold (read back through template handles), for the
base interaction and perturbed samples; the template layout; refusals."""
import tempfile

import numpy as np
import pytest


def test_multipole_basis_reproduces_oph_values():
    from nushellx_amd.synth.interaction import multipole_pn, multipole_pn_basis, random_interaction
    from nushellx_amd.synth.mspace import PF, SD, PNSpace
    for orbits, seed in ((SD, 3), (PF, 4)):
        space, base = PNSpace(orbits), random_interaction(orbits, seed=seed)
        keys, max_lmo, M = multipole_pn_basis(space, base)
        for inter in (base, base.perturbed(9, 0.1), base.unit(len(base.spe) + 5)):
            items, ml = multipole_pn(space, inter)
            assert [k for k, _ in items] == keys and ml == max_lmo
            v = np.array([x for _, x in items])
            assert np.abs(inter.parameter_vector() @ M - v).max() < 1e-13 * max(1.0, np.abs(v).max())
        assert not M[:len(base.spe)].any()  # single-particle energies do not enter the pn multipoles


def test_template_basis_reproduces_file_values():
    """theta . basis equals what a template reads from files written for theta
    (REAL*8 .mat): both sides' H(J) in the reader's partition order and the .oph
    records, for even and odd nuclei."""
    from nushellx_amd.ensemble import template_basis
    from nushellx_amd.hv import Hamiltonian
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import PNModel, SideCache, write_model
    space, base, cache = PNSpace(SD), random_interaction(SD, seed=11), SideCache()
    work = tempfile.mkdtemp(prefix="nsx_tb_")
    for (p, n), J2s in (((2, 2), [0, 4]), ((2, 3), [1, 5]), ((4, 4), [2])):
        for k, inter in enumerate((base, base.perturbed(21, 0.2))):
            m = PNModel(space, p, n, inter)
            mf = write_model(m, f"{work}/m{p}{n}_{k}", f"m{p}{n}", J2s, cache=cache, real_kind=8)
            for J2 in J2s:
                h = Hamiltonian(*mf.args(J2), device=False, template=True)
                no, hl = h.ens_layout()
                assert no == 336 and len(hl) >= 2 and {s for s, _, _ in hl} == {0, 1}
                ra, rb = m.roles()
                count = {"p": p, "n": n}
                assert all(J % 2 == count[ra if s == 0 else rb] % 2 for s, J, _ in hl)
                B = template_basis(h, space, m, cache)
                vals = h.ens_values()
                assert B.shape == (66, vals.size)
                assert np.abs(inter.parameter_vector() @ B - vals).max() < 1e-13 * np.abs(vals).max(), (p, n, J2)
                h.close()


def test_template_structure_is_a_superset():
    """A template keeps every .oph key: its terms include those of the ordinary
    handle, and its layout calls refuse ordinary handles."""
    from nushellx_amd.hv import Hamiltonian, NsxError
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import PNModel, write_model
    inter = random_interaction(SD, seed=5)
    # zero one TBME family: the ordinary handle drops the multipoles it zeroes
    for k in list(inter.tbme)[:20]:
        inter.tbme[k] = 0.0
    mf = write_model(PNModel(PNSpace(SD), 2, 2, inter), tempfile.mkdtemp(prefix="nsx_tz_"), "z22", [0], real_kind=8)
    plain = Hamiltonian(*mf.args(0), device=False)
    tmpl = Hamiltonian(*mf.args(0), device=False, template=True)
    assert plain.N == tmpl.N
    with pytest.raises(NsxError):
        plain.ens_layout()
    f_plain, _, _ = plain.work()
    f_tmpl, _, _ = tmpl.work()
    assert f_tmpl >= f_plain
