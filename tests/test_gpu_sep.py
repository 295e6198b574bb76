"""Separable pn operator (NuTRLZ) and expectation values (NuXpct) on the GPU,
against the oracle's separable restatement on the same files."""
import os

import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu

CASES = [("sd22", [0, 4]), ("sd23", [1, 5]), ("sd32", [1, 3]), ("sd44", [0, 8])]


def _err(w, ref):
    return np.abs(w - ref).max() / np.abs(ref).max()


def test_sep_hv_fp64_matches_oracle(models, dops):
    from nushellx_amd.hv import Hamiltonian
    rng = np.random.default_rng(31)
    for name, J2s in CASES:
        for J2 in J2s:
            for kind in ("cut", "exact"):
                print(f"case sep {name} 2J={J2} {kind}", flush=True)
                args = models[name].args(J2)
                ref = RefCPU(*args, fp64=True, dopfile=dops[name][kind])
                h = Hamiltonian(*args, dopfile=dops[name][kind])
                v = rng.standard_normal(h.N)
                err = _err(h.multiply(v), ref.apply(v))
                assert err < 1e-12, (name, J2, kind, err)
                h.close()


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 2e-4)])
def test_sep_hv_reduced_precision(models, dops, prec, tol):
    from nushellx_amd.hv import Hamiltonian
    rng = np.random.default_rng(32)
    for J2 in (0, 8):
        print(f"case sep sd44 2J={J2} {prec}", flush=True)
        args = models["sd44"].args(J2)
        ref = RefCPU(*args, fp64=True, dopfile=dops["sd44"]["cut"])
        h = Hamiltonian(*args, precision=prec, dopfile=dops["sd44"]["cut"])
        v = rng.standard_normal(h.N)
        assert _err(h.multiply(v), ref.apply(v)) < tol


def test_sep_batch_stacked_with_full_instances(models, dops):
    """Separable and NuTRL instances in one stacked batch."""
    from nushellx_amd.hv import Batch, DeviceBuffer, Hamiltonian, device_sync
    specs = [("sd44", 0, "cut"), ("sd44", 0, None), ("sd23", 3, "exact"), ("sd32", 1, "cut"), ("sd44", 8, "cut")]
    hams, refs = [], []
    for name, J2, kind in specs:
        args = models[name].args(J2)
        dop = dops[name][kind] if kind else None
        hams.append(Hamiltonian(*args, dopfile=dop))
        refs.append(RefCPU(*args, fp64=True, dopfile=dop))
    rng = np.random.default_rng(33)
    for cfg in ({"stacked": 1, "threads": 256, "stk_cap": 2048, "stk_work": 512},
                {"stacked": 0, "threads": 256, "stk_cap": 2048, "stk_work": 512}):
        print(f"case sep batch {cfg}", flush=True)
        b = Batch(hams, config=dict(cfg, fat_cap=5120, max_phases=6, mode=0, atomic=1))
        v = rng.standard_normal(b.N)
        din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
        b.multiply_dev(din, dout)
        device_sync()
        w = np.empty_like(v)
        dout.download(w)
        for k, r in enumerate(refs):
            o = b.offsets[k]
            assert _err(w[o:o + r.N], r.apply(v[o:o + r.N])) < 1e-12, (specs[k], cfg)
        din.free()
        dout.free()
        b.close()


def test_sep_lanczos_matches_dense(models, dops):
    from nushellx_amd.hv import Hamiltonian
    for name, J2 in (("sd44", 0), ("sd32", 3)):
        print(f"case sep lanczos {name} 2J={J2}", flush=True)
        args = models[name].args(J2)
        H = RefCPU(*args, fp64=True, dopfile=dops[name]["cut"]).dense()
        ev = np.linalg.eigvalsh(0.5 * (H + H.T))[:5]
        r = Hamiltonian(*args, dopfile=dops[name]["cut"]).lanczos(nev=5, tol=1e-8)
        assert np.abs(r["evals"] - ev).max() < 1e-7, (name, J2, r["evals"], ev)


def test_expectation_values(models, dops):
    """<v_j|O|v_i> between the lowest states of H: with O = H the matrix is
    diag(E); with another operator it equals the oracle's dots."""
    from nushellx_amd.hv import Hamiltonian
    args = models["sd44"].args(0)
    h = Hamiltonian(*args)
    r = h.lanczos(nev=5, tol=1e-9, vectors=True)
    V = r["evecs"]
    M = h.expect(V)
    assert np.abs(np.diag(M) - r["evals"]).max() < 1e-7
    assert np.abs(M - np.diag(np.diag(M))).max() < 1e-6
    for kind in ("cut", "exact"):
        print(f"case expect {kind}", flush=True)
        o = Hamiltonian(*args, dopfile=dops["sd44"][kind])
        ref = RefCPU(*args, fp64=True, dopfile=dops["sd44"][kind])
        Mo = o.expect(V)
        Mr = np.array([[ref.apply(V[i]) @ V[j] for j in range(5)] for i in range(5)])
        assert np.abs(Mo - Mr).max() < 1e-11 * np.abs(Mr).max()
        assert np.abs(Mo - Mo.T).max() < 1e-9 * np.abs(Mr).max()  # a symmetric operator
        o.close()


def test_xpct_files(models, dops, tmp_path):
    """NuXpct's file step: the solver's states written as NuVec's .xvc, the
    operator's matrix written as .ovl/.ovx (REAL*4)."""
    from nushellx_amd.hv import Hamiltonian
    from nushellx_amd.synth.fortio import read_records
    args = models["sd44"].args(4)
    h = Hamiltonian(*args)
    r = h.lanczos(nev=4, tol=1e-8, vectors=True)
    prefix = str(tmp_path / "sd44x2")
    h.write_xvc(prefix, r["evals"], r["evecs"])
    o = Hamiltonian(*args, dopfile=dops["sd44"]["cut"])
    n = o.xpct_files(prefix, "dz")
    assert n == 4
    recs = list(read_records(prefix + "dz.ovx"))
    assert int(np.frombuffer(recs[0], dtype="<i4")[0]) == 4
    Movx = np.stack([np.frombuffer(rc, dtype="<f4") for rc in recs[1:]], axis=1)  # column j = record j
    ref = RefCPU(*args, fp64=True, dopfile=dops["sd44"]["cut"])
    V32 = r["evecs"].astype(np.float32).astype(np.float64)  # the .xvc holds REAL*4 states
    Mr = np.array([[ref.apply(V32[i]) @ V32[j] for j in range(4)] for i in range(4)])
    assert np.abs(Movx - Mr).max() < 1e-5 * np.abs(Mr).max()
    txt = open(prefix + "dz.ovl").read()
    assert "Interaction/Operator file dz" in txt
    vals = [float(x) for ln in txt.splitlines()[9:] for x in ln.split()]
    assert np.allclose(vals, Movx.T.ravel(), atol=1e-4)
