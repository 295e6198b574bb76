"""Fused column path on the GPU (k_fuse + k_reduce): X = V_R Bhat^T and
Y = sum r X stay in LDS, each workgroup owns an output column for a range of
slabs, and the column partials are summed in a fixed order.  H*v against the
FP64 oracle in every precision and for several round / task sizes, bitwise
reproducibility, the path beside the large-block tiles, and the Lanczos on it."""
import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu

FUSED = {"stacked": 2, "threads": 256, "fat_cap": 5120, "max_phases": 6, "mode": 0, "atomic": 1}


def _cases(models):
    return [(n, J2) for n in ("sd22", "sd44", "sd66", "sd23", "psd33t") for J2 in models[n].J2s[:3]]


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


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-4)])
def test_fused_path_matches_refcpu(models, prec, tol):
    from nushellx_amd.hv import Batch, Hamiltonian
    cases = _cases(models)
    hams = [Hamiltonian(*models[n].args(J2), precision=prec, device=False) for n, J2 in cases]
    refs = [RefCPU(*models[n].args(J2), fp64=True) for n, J2 in cases]
    rng = np.random.default_rng(23)
    for cap, work in ((6144, 250000), (4096, 150000), (1024, 5000)):
        print(f"case fused {prec} cap={cap} work={work}", flush=True)
        b = Batch(hams, precision=prec, config=dict(FUSED, stk_cap=cap, stk_work=work))
        assert b.path().startswith("fused")
        v = rng.standard_normal(b.N)
        w = _batch_hv(b, v)
        for k, r in enumerate(refs):
            o = b.offsets[k]
            ref = r.apply(v[o:o + r.N])
            err = np.abs(w[o:o + r.N] - ref).max() / np.abs(ref).max()
            assert err < tol, (cases[k], cap, work, err)
        b.close()


def test_fused_path_is_bitwise_reproducible(models):
    from nushellx_amd.hv import Batch, Hamiltonian
    hams = [Hamiltonian(*models[n].args(J2), device=False) for n, J2 in _cases(models)]
    b = Batch(hams, config=dict(FUSED, stk_cap=4096, stk_work=150000))
    assert b.path().startswith("fused")
    v = np.random.default_rng(4).standard_normal(b.N)
    w1 = _batch_hv(b, v)
    for _ in range(3):
        assert np.array_equal(_batch_hv(b, v), w1)
    b.close()


def test_fused_path_beside_large_blocks(models, monkeypatch):
    """Items with a block side above NSX_BIG_MIN run on the matrix-core tiles
    (which add into W after the reduce has stored it); the rest on k_fuse."""
    from nushellx_amd.hv import Batch, Hamiltonian
    monkeypatch.setenv("NSX_BIG_MIN", "20")
    args = models["sd44"].args(4)
    ref = RefCPU(*args, fp64=True)
    v = np.random.default_rng(8).standard_normal(ref.N)
    w_ref = ref.apply(v)
    for prec, tol in (("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-4)):
        b = Batch([Hamiltonian(*args, precision=prec, device=False)], precision=prec,
                  config=dict(FUSED, stk_cap=4096, stk_work=150000))
        assert b.path().startswith("fused")
        err = np.abs(_batch_hv(b, v) - w_ref).max() / np.abs(w_ref).max()
        assert err < tol, (prec, err)
        b.close()


def test_fused_path_lanczos_matches_dense(models):
    from nushellx_amd.hv import Batch, Hamiltonian
    cases = [("sd44", 0), ("sd44", 8), ("sd22", 2), ("psd33t", 2)]
    hams = [Hamiltonian(*models[n].args(J2), device=False) for n, J2 in cases]
    b = Batch(hams, config=dict(FUSED, stk_cap=4096, stk_work=150000))
    r = b.lanczos(nev=5, tol=1e-8)
    assert (r["converged_at"] >= 0).all()
    for k, (n, J2) in enumerate(cases):
        ref = RefCPU(*models[n].args(J2), fp64=True)
        _, keep = ref.trunc()
        idx = np.flatnonzero(keep)
        H = ref.dense()[np.ix_(idx, idx)]
        ev = np.linalg.eigvalsh(0.5 * (H + H.T))[:5]
        assert np.abs(r["evals"][k] - ev).max() < 1e-7, (n, J2)
    b.close()
