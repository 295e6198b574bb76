"""N-hbar-omega truncated bases on the GPU: H*v against the truncated oracle,
and the lowest eigenvalues against the full operator restricted to the window
states (P H P, from the untruncated oracle)."""
import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu


def test_truncated_hv_matches_oracle(models):
    from nushellx_amd.hv import Hamiltonian
    rng = np.random.default_rng(41)
    for name in ("psd33t", "psd33w8"):
        for J2 in models[name].J2s:
            print(f"case trunc hv {name} 2J={J2}", flush=True)
            args = models[name].args(J2)
            ref = RefCPU(*args, fp64=True)
            h = Hamiltonian(*args)
            t, m = h.trunc()
            assert t and np.array_equal(m, ref.trunc()[1])
            v = rng.standard_normal(h.N)
            w_ref = ref.apply(v)
            assert np.abs(h.multiply(v) - w_ref).max() <= 1e-12 * np.abs(w_ref).max()
            h.close()


def test_truncated_lanczos_matches_window_spectrum(models):
    from nushellx_amd.hv import Batch, Hamiltonian
    evs, hams = [], []
    for J2 in models["psd33t"].J2s:
        print(f"case trunc lanczos 2J={J2}", flush=True)
        H = RefCPU(*models["psd33"].args(J2), fp64=True).dense()
        t, m = RefCPU(*models["psd33t"].args(J2)).trunc()
        idx = np.where(m)[0]
        Hw = H[np.ix_(idx, idx)]
        ev = np.linalg.eigvalsh(0.5 * (Hw + Hw.T))[:4]
        h = Hamiltonian(*models["psd33t"].args(J2))
        r = h.lanczos(nev=4, tol=1e-8, max_basis=min(40, len(idx)), keep=min(10, len(idx) - 2))
        assert np.abs(r["evals"] - ev).max() < 1e-7, (J2, r["evals"], ev)
        evs.append(ev)
        hams.append(h)
    # truncated and untruncated instances in one batch
    full0 = Hamiltonian(*models["psd33"].args(0))
    rf = RefCPU(*models["psd33"].args(0), fp64=True)
    lc = np.where(rf.lcoef())[0]  # positive parity only (tparity = +1)
    Hf = rf.dense()[np.ix_(lc, lc)]
    ev_full = np.linalg.eigvalsh(0.5 * (Hf + Hf.T))[:4]
    rb = Batch(hams + [full0]).lanczos(nev=4, tol=1e-8, max_basis=24, keep=8)
    for k, ev in enumerate(evs + [ev_full]):
        assert np.abs(rb["evals"][k] - ev).max() < 1e-7, (k, rb["evals"][k], ev)
