"""Ensembles: interaction samples x nuclei x J as host-only instances planned
as one batch.  Every instance's lowest 5 eigenvalues equal its own
single-instance device solve."""
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_ensemble_batch_matches_single_instances():
    from nushellx_amd.ensemble import build_instances, solve_ensemble, write_ensemble
    from nushellx_amd.hv import Hamiltonian, NsxError
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    J2s = [0, 4]
    nuclei = [(2, 2), (4, 4)]
    base = random_interaction(SD, seed=2024)
    entries = write_ensemble(PNSpace(SD), nuclei, J2s, base, seeds=[11, 12, 13, 14],
                             workdir=tempfile.mkdtemp(prefix="nsx_ens_"))
    hams = build_instances(entries, J2s, threads=8)
    assert len(hams) == 4 * 2 * 2
    with pytest.raises(NsxError):
        hams[0].multiply(np.zeros(hams[0].N))
    n_tune = len(nuclei) * len(J2s)
    r = solve_ensemble(hams, nev=5, tol=1e-8, tune_on=build_instances(entries[:len(nuclei)], J2s))
    assert (r["converged_at"] >= 0).all()
    jobs = [(mf, J2) for (_, _, mf) in entries for J2 in J2s]
    for k, (mf, J2) in enumerate(jobs):
        single = Hamiltonian(*mf.args(J2)).lanczos(nev=5, tol=1e-8)
        assert np.abs(r["evals"][k] - single["evals"]).max() < 1e-7, (k, J2)
    # samples differ: the same nucleus and J gives different spectra per sample
    assert np.abs(r["evals"][0] - r["evals"][n_tune]).max() > 1e-3
