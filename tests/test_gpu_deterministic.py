"""Deterministic H*v (NSX_DETERMINISTIC=1): for the small blocks the fused
matrix-core path storing per-task partial tiles that one workgroup per W tile
sums in task order (or, where a tuning result asks for it, the fused column
path, column partials summed in plan order), and one stage-2 workgroup per
large-block W tile (a single addition onto the stored like-particle value).
Repeated H*v are bitwise identical, and equal the FP64 oracle, on a model with
both kinds of blocks (pf 4p4n, sides up to 99) and on an sd batch."""
import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("bf16x3", 2e-4)])
def test_deterministic_pf44(pf44, prec, tol, monkeypatch):
    from nushellx_amd.hv import Hamiltonian
    monkeypatch.setenv("NSX_DETERMINISTIC", "1")
    args = pf44.args(0)
    ref = RefCPU(*args, fp64=True)
    v = np.random.default_rng(6).standard_normal(ref.N)
    w_ref = ref.apply(v)
    print(f"case deterministic pf44 {prec}", flush=True)
    h = Hamiltonian(*args, precision=prec)
    w = h.multiply(v)
    assert np.abs(w - w_ref).max() / np.abs(w_ref).max() < tol
    for _ in range(3):
        assert np.array_equal(h.multiply(v), w)
    h.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
@pytest.mark.parametrize("stacked,path", [(None, "fused-mfma +like deterministic"), (2, "fused deterministic")])
def test_deterministic_sd_batch(models, monkeypatch, stacked, path, prec, tol):
    from nushellx_amd.hv import Batch, DeviceBuffer, Hamiltonian, device_sync
    monkeypatch.setenv("NSX_DETERMINISTIC", "1")
    cases = [(n, J2) for n in ("sd22", "sd44", "sd66") for J2 in models[n].J2s[:3]]
    print(f"case deterministic sd batch {stacked} {prec}", flush=True)
    hams = [Hamiltonian(*models[n].args(J2), precision=prec, device=False) for n, J2 in cases]
    cfg = None if stacked is None else {"fat_cap": 5120, "max_phases": 6, "threads": 256, "mode": 0, "atomic": 1,
                                        "stacked": stacked, "deterministic": 1}
    b = Batch(hams, precision=prec, config=cfg)
    assert b.path() == path
    v = np.random.default_rng(7).standard_normal(b.N)
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    outs = []
    for _ in range(3):
        b.multiply_dev(din, dout)
        device_sync()
        w = np.empty_like(v)
        dout.download(w)
        outs.append(w)
    assert all(np.array_equal(o, outs[0]) for o in outs[1:])
    for k, (n, J2) in enumerate(cases):
        r = RefCPU(*models[n].args(J2), fp64=True)
        o = b.offsets[k]
        ref = r.apply(v[o:o + r.N])
        assert np.abs(outs[0][o:o + r.N] - ref).max() <= tol * np.abs(ref).max()
    b.close()
