"""NuVec's program step on the GPU (nsx_nuvec): the .abz + .lnz/.trl that the
solver writes, or that a Lanczos run with a genuine tridiagonal / thick-restart
history leaves, turned into .xvc states, .xva averages, .lev and .lp.  Checked
against the oracle's restatement of NuVec (oracle/nuvec.py): the states (REAL*4),
the energies, the per-partition percentages, the average shell occupations
and the parity shares."""
import os
import tempfile

import numpy as np
import pytest

from oracle import nuvec as onv
from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu


def _xvc(prefix):
    recs = onv.read_records(prefix + ".xvc")
    n, nlev = np.frombuffer(recs[0][:8], dtype="<i4")
    E = np.array([np.frombuffer(r[:4], dtype="<f4")[0] for r in recs[1:1 + nlev]], dtype=np.float64)
    X = np.array([np.frombuffer(r[4:], dtype="<f4") for r in recs[1:1 + nlev]], dtype=np.float64)
    return int(n), E, X


def _check(h, args, prefix, nlev_expect):
    nlev = h.nuvec(prefix)
    assert nlev == nlev_expect
    w, S = onv.nuvec_states_any(prefix)
    n, E, X = _xvc(prefix)
    assert n == h.N and len(E) == nlev
    assert np.abs(E - w).max() <= 1e-5 * max(1.0, np.abs(w).max())
    for q in range(nlev):  # same eigenvector sign convention need not hold: compare up to sign
        s = np.sign(X[q] @ S[q])
        assert np.abs(X[q] - s * S[q]).max() < 2e-6 * max(1.0, np.abs(S[q]).max())
    ana = onv.analyse(X, args[0], args[1], args[2], onv.read_abz(prefix)["bas"])
    ns = len(onv.read_nus_shells(args[0]))
    xva = onv.read_xva(prefix, ns, ana[0]["spartpc"].shape)
    assert len(xva) == nlev
    for q in range(nlev):
        assert abs(xva[q]["energy"] - E[q]) < 1e-4
        assert np.abs(xva[q]["spartpc"] - ana[q]["spartpc"]).max() < 1e-3
        assert abs(xva[q]["spartpc"].sum() - 100.0) < 1e-2
        assert np.abs(xva[q]["ava"] - ana[q]["ava"]).max() < 1e-4 * max(1.0, ana[q]["ava"].max())
        assert np.abs(xva[q]["avb"] - ana[q]["avb"]).max() < 1e-4 * max(1.0, ana[q]["avb"].max())
    lev = onv.read_lev(prefix)
    assert [round(e, 3) for e, _ in lev] == [round(float(e), 3) for e in E]
    return E, X, lev, ana


def test_nuvec_on_solver_states(models):
    from nushellx_amd.hv import Hamiltonian
    args = models["sd44"].args(8)
    h = Hamiltonian(*args)
    r = h.lanczos(nev=6, tol=1e-8, vectors=True)
    d = tempfile.mkdtemp(prefix="nsx_nuvec_")
    for layout in ("trl", "lnz"):
        prefix = os.path.join(d, layout, "sd444")
        os.makedirs(os.path.dirname(prefix))
        print(f"case nuvec {layout}", flush=True)
        if layout == "trl":
            h.write_trl(prefix, r["evals"], r["evecs"], no_level=5)
        else:
            h.write_vectors(prefix, r["evals"], r["evecs"], no_level=5)
        E, X, lev, ana = _check(h, args, prefix, 5)
        assert np.abs(E - r["evals"][:5]).max() < 1e-4
        # eigenstates of a positive-parity sd model: all amplitude in positive-parity partition pairs
        assert all(mark == "+" for _, mark in lev)
        assert all(a["nparamp"] < 1e-12 * a["tampl"] for a in ana)
        # the states are eigenvectors of the oracle's H (REAL*4 storage)
        ref = RefCPU(*args, fp64=True)
        for q in range(5):
            res = ref.apply(X[q]) - E[q] * X[q]
            assert np.linalg.norm(res) < 1e-3 * abs(E[q])
    h.close()


def test_nuvec_on_lanczos_histories(models):
    """Files with a genuine projected matrix: a tridiagonal (alpha, beta) for
    the .lnz layout and for a thick-restart run without restart (count = 1),
    and an arrowhead-plus-tridiagonal matrix after a restart (count = 2)."""
    from nushellx_amd.hv import Hamiltonian
    from nushellx_amd.synth.fortio import RecordWriter
    args = models["sd44"].args(4)
    h = Hamiltonian(*args)
    rng = np.random.default_rng(12)
    n = 12
    lc = h.lcoef()
    Q = np.zeros((h.N, n))
    Q[lc], _ = np.linalg.qr(rng.standard_normal((int(lc.sum()), n)))
    alpha = rng.uniform(-40, -10, n)
    beta = np.abs(rng.normal(2.0, 1.0, n))
    bas = np.frombuffer(b"".join(bytes([a, b, 0, 0, 0, 0, 0, 0]) for a, b, _, _ in h.blocks()), dtype=np.uint8)
    d = tempfile.mkdtemp(prefix="nsx_nuvec_h_")
    for kind in ("lnz", "trl1", "trl2"):
        prefix = os.path.join(d, kind, "sd442")
        os.makedirs(os.path.dirname(prefix))
        print(f"case nuvec history {kind}", flush=True)
        w = RecordWriter(prefix + ".abz")
        if kind == "lnz":  # Lnzs: beta(i) couples i-1 and i
            w.write(n, alpha.astype(np.float32), np.concatenate([[0.0], beta[:-1]]).astype(np.float32))
            w.write(n, 5, np.float32(0.0), h.N, h.basd)
            w.write(bas)
        else:
            M = np.zeros((n, n))
            count = 1 if kind == "trl1" else 2
            if count == 2:  # restart: arrowhead block of 4 kept values, then the tridiagonal tail
                M[:n, :n] = np.diag(alpha)
                M[:4, 4] = M[4, :4] = beta[:4]
                for i in range(4, n - 1):
                    M[i, i + 1] = M[i + 1, i] = beta[i]
            w.write(n, alpha.astype(np.float32), beta.astype(np.float32), M.astype(np.float32))
            w.write(n, 5, np.float32(0.0), h.N, h.basd, count, 4, 0, 80, 3, 300)
            w.write(bas)
            w.write(np.zeros(300, np.float32), np.float32(0.0), np.arange(1, 301, dtype=np.int32))
        w.close()
        Q.T.astype(np.float32).tofile(prefix + (".lnz" if kind == "lnz" else ".trl"))
        _check(h, args, prefix, 5)
    h.close()
