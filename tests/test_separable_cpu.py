"""The separable (Dufour-Zuker) pn operator of NuTRLZ / NuXpct, on the CPU.

Pins, in order:
* the NuDiop restatement: with no cut the eigen-terms rebuild the pn multipole
  interaction of the .oph exactly (sum_n e_n u_n u_n^T = V);
* the oracle's separable H*v: with all terms kept it equals the oracle's full
  NuTRL H*v (independently pinned), up to the REAL*4 rounding of the
  eigenvalues the reader imposes; a truncated file does not;
* the product's separable model, run through its device plan by the CPU
  emulator, against the oracle's separable H*v (both a/b role assignments);
* list-directed reading: a reflowed file gives the same operator.
"""
import os

import numpy as np
import pytest

from nushellx_amd.synth.diop import diop_sections, read_oph, separable_matrix, write_dop
from oracle.refcpu import RefCPU
from tests.emulator import emu_hv_sep


def test_diop_rebuilds_interaction(models):
    for name in ("sd22", "sd32"):
        items, max_lmo = read_oph(models[name].intfile + ".oph")
        secs = diop_sections(items, max_lmo)
        assert len(secs) == 2 * (max_lmo // 2 + 1)
        worst, used = 0.0, 0
        for lm in range(0, max_lmo + 1, 2):
            for pt in (-1, 1):
                S = separable_matrix(secs, lm, pt)
                for key, me, pl, pr in items:
                    if key[4] != lm or pl != pt or pr != pt or 100 * key[0] + key[1] > 100 * key[2] + key[3]:
                        continue
                    worst = max(worst, abs(S.get(((key[0], key[1]), (key[2], key[3])), 0.0) - me))
                    used += 1
        assert used > 100
        assert worst < 1e-12, (name, worst)
        # every eigenvector is normalised and the pair parities of a section agree
        for lm, pt, terms in secs:
            for e, lines in terms:
                assert abs(sum(l[4] ** 2 for l in lines) - 1.0) < 1e-12


def test_diop_cut_drops_small_terms(models):
    items, max_lmo = read_oph(models["sd22"].intfile + ".oph")
    full = diop_sections(items, max_lmo)
    cut = diop_sections(items, max_lmo, cut=0.3)
    for (lm, pt, tf), (_, _, tc) in zip(full, cut):
        keep = [e for e, _ in tf if lm == 0 or abs(e) / np.sqrt(lm + 1) > 0.3]
        assert [e for e, _ in tc] == keep


def test_oracle_separable_equals_full(models, dops):
    rng = np.random.default_rng(3)
    for name, J2 in (("sd22", 4), ("sd23", 3), ("sd32", 1), ("sd44", 0)):
        args = models[name].args(J2)
        full = RefCPU(*args, fp64=True)
        v = rng.standard_normal(full.N)
        w = full.apply(v)
        sc = np.abs(w).max()
        for limits in (False, True):
            sep = RefCPU(*args, fp64=True, dopfile=dops[name]["exact"], limits=limits)
            err = np.abs(sep.apply(v) - w).max() / sc
            # eigenvalues are REAL*4 in the reader: relative rounding 2^-24 per term
            assert err < 2e-7, (name, J2, limits, err)
        cut = RefCPU(*args, fp64=True, dopfile=dops[name]["cut"])
        assert np.abs(cut.apply(v) - w).max() / sc > 1e-3
        f7 = RefCPU(*args, fp64=True, dopfile=dops[name]["ref"])
        assert np.abs(f7.apply(v) - w).max() / sc < 1e-6


def test_product_separable_plan_matches_oracle(models, dops):
    rng = np.random.default_rng(5)
    for name, J2s in (("sd22", [0, 4]), ("sd23", [1, 5]), ("sd32", [1, 3]), ("sd44", [8])):
        for J2 in J2s:
            args = models[name].args(J2)
            for kind in ("exact", "cut", "ref"):
                ref = RefCPU(*args, fp64=True, dopfile=dops[name][kind])
                v = rng.standard_normal(ref.N)
                w_ref = ref.apply(v)
                w, _ = emu_hv_sep(args, dops[name][kind], v)
                err = np.abs(w - w_ref).max() / np.abs(w_ref).max()
                assert err < 1e-12, (name, J2, kind, err)


def test_dop_list_directed_reflow(models, dops, tmp_path):
    """Component lines split over two lines and trailing values after each READ's
    items parse to the same operator (Fortran list-directed rules)."""
    name, J2 = "sd22", 4
    src = dops[name]["cut"] + ".dop"
    lines = open(src).read().splitlines()
    out = []
    for ln in lines:
        t = ln.split()
        if len(t) == 5:                      # component: split after the shells
            out += [" ".join(t[:2]), " ".join(t[2:])]
        elif len(t) == 2:                    # section header: trailing junk
            out.append(ln + "  7  9")
        else:
            out.append(ln + "  123")         # counts and eigenvalues: junk after the item
    dst = str(tmp_path / "reflow")
    open(dst + ".dop", "w").write("\n".join(out) + "\n")
    args = models[name].args(J2)
    v = np.random.default_rng(1).standard_normal(RefCPU(*args).N)
    w0 = RefCPU(*args, dopfile=dops[name]["cut"]).apply(v)
    # the oracle's threaded reduction order varies between runs: rounding-level agreement
    assert np.abs(RefCPU(*args, dopfile=dst).apply(v) - w0).max() <= 1e-14 * np.abs(w0).max()
    w1, _ = emu_hv_sep(args, dst, v)
    w2, _ = emu_hv_sep(args, dops[name]["cut"], v)
    assert np.array_equal(w1, w2)


def test_missing_dop_is_an_error(models):
    from tests.emulator import lib
    args = models["sd22"].args(0)
    with pytest.raises(RuntimeError, match="cannot open"):
        emu_hv_sep(args, "/nonexistent/op", np.zeros(RefCPU(*args).N))
    with pytest.raises(RuntimeError):
        RefCPU(*args, dopfile="/nonexistent/op")
