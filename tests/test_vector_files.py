"""Lanczos-vector files (.lnz/.abz) in the reference layout, read back by the
oracle's restatement of NuVec: the states NuVec would build are the written
Ritz vectors (REAL*4) with their energies, and the header carries the basis
size and block keys.  Host-only handle: no GPU needed."""
import os
import tempfile

import numpy as np

from oracle.nuvec import nuvec_states, read_abz_lnz


def test_write_vectors_round_trip_through_nuvec(models):
    from nushellx_amd.hv import Hamiltonian
    mf = models["sd22"]
    h = Hamiltonian(*mf.args(4), device=False)
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((h.N, 4)))
    vecs = Q.T
    evals = np.array([-12.5, -11.0, -9.25, -3.0])
    prefix = os.path.join(tempfile.mkdtemp(), "sd22" + "2")
    h.write_vectors(prefix, evals, vecs, no_level=3, tt=1.5)
    abz = read_abz_lnz(prefix)
    assert abz["no_state"] == h.N and abz["isave"] == 4 and abz["no_level"] == 3 and abz["tt"] == 1.5
    assert abz["basd"] == h.basd
    blocks = h.blocks()
    assert [(k[0], k[1]) for k in abz["bas"]] == [(b[0], b[1]) for b in blocks]
    assert all(k[2:] == bytes(6) for k in abz["bas"])
    assert os.path.getsize(prefix + ".lnz") == 4 * 4 * h.N  # four records of N REAL*4, no markers
    E, X = nuvec_states(prefix)
    assert np.allclose(E, evals[:3], atol=1e-6)
    for q in range(3):
        assert np.abs(np.abs(X[q] @ vecs[q]) - 1.0) < 1e-6
        assert np.abs(X[q] - vecs[q]).max() < 1e-6  # diagonal tridiagonal: no sign flip, REAL*4 rounding


def test_write_trl_round_trip_through_nuvec(models):
    """The thick-restart layout (.trl + four-record .abz) read by the oracle's
    restatement of NuVec's TRL branch: count = 2, the projected matrix, the
    fourth record, and the states it rebuilds."""
    from oracle.nuvec import nuvec_states_any, read_abz
    from nushellx_amd.hv import Hamiltonian
    mf = models["sd44"]
    h = Hamiltonian(*mf.args(2), device=False)
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((h.N, 6)))
    vecs = Q.T
    evals = np.array([-30.5, -29.0, -28.25, -27.0, -26.5, -25.0])
    prefix = os.path.join(tempfile.mkdtemp(), "sd44" + "1")
    h.write_trl(prefix, evals, vecs, no_level=5, tt=2.0)
    assert not os.path.exists(prefix + ".lnz")
    abz = read_abz(prefix)
    assert abz["trl"] and abz["count"] == 2 and abz["isave"] == 6 and abz["no_level"] == 5
    assert abz["no_state"] == h.N and abz["basd"] == h.basd
    assert np.allclose(abz["matrix"], np.diag(evals), atol=1e-5)
    assert len(abz["rec4"]) == 4 * 300 + 4 + 4 * 300
    assert os.path.getsize(prefix + ".trl") == 6 * 4 * h.N
    E, X = nuvec_states_any(prefix)
    assert np.allclose(E, evals[:5], atol=1e-5)
    for q in range(5):
        assert np.abs(np.abs(X[q] @ vecs[q]) - 1.0) < 1e-6
