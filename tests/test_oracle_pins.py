"""Pins of the oracle and of the generator against closed forms, library
routines, the reference's own documented examples, and an independent M-scheme
shell-model code.  CPU only."""
import json
import os

import numpy as np
import pytest

from nushellx_amd.synth import angmom
from nushellx_amd.synth.interaction import random_interaction
from nushellx_amd.synth.mspace import PF, SD, PNSpace
from nushellx_amd.synth.side import SideBasis, comp_next_all
from nushellx_amd.synth.writer import PNModel, SideCache, write_model
from oracle import refcpu
from oracle.mscheme import MSchemeOracle, read_int_file
from oracle.refcpu import RefCPU
from tests import emulator

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "angmom.json")))


def test_sixj_golden():
    """6j: oracle (binomial form), product (Racah sum) and generator vs sympy."""
    for a, b, c, d, e, f, v in GOLD["sixj"]:
        assert abs(refcpu.sixj(a, b, c, d, e, f) - v) < 1e-12
        assert abs(emulator.sixj(a, b, c, d, e, f) - v) < 1e-12
        assert abs(angmom.sixj(a, b, c, d, e, f) - v) < 1e-12


def test_cg_golden():
    for j1, j2, j3, m1, m2, m3, v in GOLD["cg"]:
        assert abs(angmom.cg(j1, j2, j3, m1, m2, m3) - v) < 1e-13


def test_racah_symmetry():
    """W(abcd;ef) = W(badc;ef) = W(cdab;ef) (Racah symmetries)."""
    for (a, b, c, d, e, f) in [(5, 3, 1, 5, 4, 6), (7, 5, 3, 1, 6, 4), (8, 4, 6, 2, 6, 4)]:
        w = angmom.racah(a, b, c, d, e, f)
        assert abs(w - angmom.racah(b, a, d, c, e, f)) < 1e-14
        assert abs(w - angmom.racah(c, d, a, b, e, f)) < 1e-14


def test_comp_next_documented_example():
    """comp_next (src/NuBasis/NuBasis.f90:271-354): the set of compositions of 6
    into 3 parts equals the 28 listed in its comment (:285-290).  The comment
    lists them in reverse-lexicographic order; the code (NEXCOM with SAVEd it/ih)
    produces a different order, which is the one that defines NuShellX's
    partition numbering.  The order is pinned by a hand trace of that code:
    from (6,0,0): it=6>1 -> ih=1, it=a(1)=6, a(1)=5, a(2)=1 -> (5,1,0); then
    it=6>1 -> ih=1, it=5, a(1)=4, a(2)=2 -> (4,2,0); ... (1,5,0) -> ih=1, it=1,
    a=(0,6,0); it=1 -> ih=2, it=a(2)=6, a(2)=0, a(1)=5, a(3)=1 -> (5,0,1)."""
    want = [(6, 0, 0), (5, 1, 0), (5, 0, 1), (4, 2, 0), (4, 1, 1), (4, 0, 2), (3, 3, 0), (3, 2, 1),
            (3, 1, 2), (3, 0, 3), (2, 4, 0), (2, 3, 1), (2, 2, 2), (2, 1, 3), (2, 0, 4), (1, 5, 0),
            (1, 4, 1), (1, 3, 2), (1, 2, 3), (1, 1, 4), (1, 0, 5), (0, 6, 0), (0, 5, 1), (0, 4, 2),
            (0, 3, 3), (0, 2, 4), (0, 1, 5), (0, 0, 6)]
    got = comp_next_all(6, 3)
    assert sorted(got) == sorted(want) and len(got) == 28
    assert got[:9] == [(6, 0, 0), (5, 1, 0), (4, 2, 0), (3, 3, 0), (2, 4, 0), (1, 5, 0), (0, 6, 0),
                       (5, 0, 1), (4, 1, 1)]
    assert got[-1] == (0, 0, 6)


@pytest.mark.parametrize("orb,n,dims", [
    (SD, 4, {0: 9, 2: 12, 4: 21, 6: 15, 8: 15, 10: 6, 12: 3}),
    (SD, 6, {0: 14, 2: 19, 4: 33, 6: 29, 8: 26, 10: 12, 12: 8, 14: 1}),
    (SD, 5, {1: 15, 3: 24, 5: 29, 7: 23, 9: 17, 11: 8, 13: 3}),
])
def test_side_dimensions(orb, n, dims):
    s = SideBasis(PNSpace(orb), "p", n)
    assert {J2: s.dJ(J2) for J2 in s.J2s} == dims


@pytest.mark.parametrize("orb,n,N0", [(SD, 2, 46), (SD, 4, 1161), (SD, 6, 3372), (PF, 4, 41355)])
def test_pn_J0_dimension_textbook(orb, n, N0):
    """pn J=0 dimension = sum_J d(J)^2: 20Ne 46, 24Mg 1161, 28Si 3372, 48Cr 41355."""
    s = SideBasis(PNSpace(orb), "p", n)
    assert sum(s.dJ(J2) ** 2 for J2 in s.J2s) == N0


def test_nJT_counting_identity():
    """nJT(p, J) = #(M=J) - #(M=J+1) and sum_J (2J+1) d(J) = C(12, n) (sd, 4 particles)."""
    s = SideBasis(PNSpace(SD), "n", 4)
    total = sum((J2 + 1) * s.dJ(J2) for J2 in s.J2s)
    assert total == 495


def test_number_operator_density_closed_form():
    """[a+ a~]^0 = N_s / sqrt(2j+1), so the lambda=0, s=s' density between
    states of one partition is n_s(p) sqrt(2J+1)/sqrt(2j_s+1) times identity."""
    s = SideBasis(PNSpace(SD), "p", 4)
    dens = s.densities(10)
    for J2 in s.J2s:
        off = np.concatenate([[0], np.cumsum(s.nJT[:, s.jidx(J2)])])
        for li, sh in enumerate(s.shells):
            D = dens.get((J2, J2, sh, sh, 0))
            if D is None:
                continue
            want = np.zeros_like(D)
            for p, occ in enumerate(s.partitions):
                a, b = off[p], off[p + 1]
                want[a:b, a:b] = np.eye(b - a) * occ[li] * np.sqrt(J2 + 1) / np.sqrt(s.j2[li] + 1)
            assert np.abs(D - want).max() < 1e-12


def test_jbasis_orthonormal_and_J_eigenstates():
    s = SideBasis(PNSpace(SD), "n", 3)
    for J2 in s.J2s:
        for B in s.phi(J2):
            if B.shape[1]:
                assert np.abs(B.T @ B - np.eye(B.shape[1])).max() < 1e-12


@pytest.fixture(scope="module")
def real8_models(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("r8"))
    inter = random_interaction(SD, seed=99)
    cache = SideCache()
    sd = PNSpace(SD)
    return {
        "22": write_model(PNModel(sd, 2, 2, inter), d + "/a", "r22", list(range(0, 17, 2)), cache=cache, real_kind=8),
        "12": write_model(PNModel(sd, 1, 2, inter), d + "/b", "r12", list(range(1, 14, 2)), cache=cache, real_kind=8),
    }


@pytest.mark.parametrize("key,np_,nn,M2", [("22", 2, 2, 0), ("12", 1, 2, 1)])
def test_jscheme_union_equals_mscheme(real8_models, key, np_, nn, M2):
    """Independent pin of the whole path on REAL*8 operator files: the union over
    J of the J-scheme spectra (oracle H*v, dense) equals the M-scheme M=M0 spectrum."""
    mf = real8_models[key]
    ev = []
    for J2 in mf.J2s:
        r = RefCPU(*mf.args(J2), fp64=True)
        if r.N == 0:
            continue
        H = r.dense()
        assert np.abs(H - H.T).max() < 1e-10
        ev.extend(np.linalg.eigvalsh(0.5 * (H + H.T)))
    spe, tb = read_int_file(mf.intfile + ".int")
    m = MSchemeOracle([5, 1, 3], spe, tb, np_, nn)
    em = np.linalg.eigvalsh(m.hamiltonian(M2).toarray())
    ev = np.sort(ev)
    assert len(ev) == len(em)
    assert np.abs(ev - em).max() < 1e-9


def test_refcpu_real4_build_and_thresholds(models):
    """The REAL*4 build with the 1e-7 product thresholds stays within 1e-5 of FP64."""
    args = models["sd22"].args(0)
    h64 = RefCPU(*args, fp64=True).dense()
    h32 = RefCPU(*args, fp64=False, thresholds=True).dense()
    e64 = np.linalg.eigvalsh(0.5 * (h64 + h64.T))
    e32 = np.linalg.eigvalsh(0.5 * (h32 + h32.T))
    assert np.abs(e64 - e32).max() < 1e-4
    assert np.abs(h64 - h32).max() > 0  # the builds really differ


def test_tfind_float_bisection_finds_every_key():
    """The reference's float-bisection key search (src/NuCommonX/NuCommonX.f90:147-183)
    returns every present key for array sizes 1..1500 and sampled sizes to 1e5,
    so the exact search used by the product is equivalent."""
    rng = np.random.default_rng(0)
    for n in list(range(1, 1500, 7)) + [2048, 4097, 65536, 100000]:
        keys = np.unique(rng.integers(0, 2 ** 62, size=2 * n, dtype=np.int64))[:n]
        kb = b"".join(int(k).to_bytes(8, "big") for k in np.sort(keys))
        n = len(keys)
        for i in (range(n) if n < 2000 else rng.integers(0, n, 300)):
            assert refcpu.tfind(kb, kb[8 * i:8 * i + 8]) == i + 1
        assert refcpu.tfind(kb, b"\xff" * 8) == 0
