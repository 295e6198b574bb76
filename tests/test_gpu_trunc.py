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


def _batch_hv(b, v):
    from nushellx_amd.hv import DeviceBuffer, device_sync
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    b.multiply_dev(din, dout)
    device_sync()
    w = np.empty_like(v)
    dout.download(w)
    din.free()
    dout.free()
    return w


@pytest.mark.parametrize("config", [None, {"stacked": 1, "threads": 256, "fat_cap": 5120, "max_phases": 6,
                                           "mode": 0, "atomic": 1},
                                    {"stacked": 2, "threads": 256, "fat_cap": 5120, "max_phases": 6, "mode": 0,
                                     "atomic": 1, "stk_cap": 4096, "stk_work": 150000},
                                    {"stacked": 3, "threads": 256, "fat_cap": 5120, "max_phases": 6, "mode": 0,
                                     "atomic": 1, "stk_cap": 8, "stk_work": 4},
                                    {"stacked": 4, "threads": 256, "fat_cap": 5120, "max_phases": 6, "mode": 0,
                                     "atomic": 1, "stk_cap": 1, "stk_work": 4}])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-4)])
def test_truncated_pn_input_is_masked(models, config, prec, tol):
    """A vector nonzero outside the window: the pn part sees only its window
    states, the like-particle part all of them (the oracle restates NuTRL's
    element masks); truncated and full instances in one batch, every path."""
    from nushellx_amd.hv import Batch, Hamiltonian
    cases = [("psd33t", 0), ("psd33", 0), ("psd33t", 2), ("psd33w8", 0), ("psd33t", 4)]
    print(f"case trunc mask {prec} {config and config['stacked']}", flush=True)
    hams = [Hamiltonian(*models[n].args(J2), precision=prec, device=False) for n, J2 in cases]
    b = Batch(hams, precision=prec, config=config)
    if config and config["stacked"] == 4:  # the like-particle part stays out of the fused path here
        assert b.path() == "fused-mfma"
    v = np.random.default_rng(43).standard_normal(b.N)
    w = _batch_hv(b, v)
    for k, (n, J2) in enumerate(cases):
        ref = RefCPU(*models[n].args(J2), fp64=True)
        o = b.offsets[k]
        w_ref = ref.apply(v[o:o + ref.N])
        assert np.abs(w[o:o + ref.N] - w_ref).max() <= tol * np.abs(w_ref).max(), (n, J2)
    b.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("bf16x3", 2e-4)])
def test_truncated_pn_input_is_masked_on_large_blocks(models, monkeypatch, prec, tol):
    """The same with the blocks above side 12 on the matrix-core tiles."""
    from nushellx_amd.hv import Batch, Hamiltonian
    monkeypatch.setenv("NSX_BIG_MIN", "12")
    cases = [("psd33t", 2), ("psd33", 2), ("psd33t", 4)]
    print(f"case trunc mask large blocks {prec}", flush=True)
    hams = [Hamiltonian(*models[n].args(J2), precision=prec, device=False) for n, J2 in cases]
    b = Batch(hams, precision=prec)
    v = np.random.default_rng(47).standard_normal(b.N)
    w = _batch_hv(b, v)
    for k, (n, J2) in enumerate(cases):
        ref = RefCPU(*models[n].args(J2), fp64=True)
        o = b.offsets[k]
        w_ref = ref.apply(v[o:o + ref.N])
        assert np.abs(w[o:o + ref.N] - w_ref).max() <= tol * np.abs(w_ref).max(), (n, J2)
    b.close()
