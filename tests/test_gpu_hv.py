"""GPU H*v parity: libnsx.so (HIP, gfx950) against the FP64 restatement of
NuTRL's multiplyx/multiplya/multiplyb (oracle/refcpu) on the same files."""
import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative, FP64 vs FP64 with a different summation order


def _cases(models):
    for name, mf in models.items():
        for J2 in mf.J2s:
            yield name, J2, mf.args(J2)


def test_hv_fp64_matches_refcpu(models):
    from nushellx_amd.hv import Hamiltonian
    rng = np.random.default_rng(7)
    n = 0
    for name, J2, args in _cases(models):
        print(f"case {name} 2J={J2}", flush=True)
        ref = RefCPU(*args, fp64=True)
        h = Hamiltonian(*args)
        assert h.N == ref.N
        assert [(a, b, p, d) for (a, b, p, d) in h.blocks()] == ref.blocks()
        assert np.array_equal(h.lcoef(), ref.lcoef())
        for _ in range(2):
            v = rng.standard_normal(h.N)
            w_ref = ref.apply(v)
            w = h.multiply(v)
            err = np.abs(w - w_ref).max() / max(np.abs(w_ref).max(), 1e-300)
            assert err < TOL, (name, J2, err)
        n += 1
    assert n >= 20


def test_hv_dev_and_f32_entry_points(models):
    from nushellx_amd.hv import DeviceBuffer, Hamiltonian, device_sync
    args = models["sd44"].args(4)
    h = Hamiltonian(*args)
    ref = RefCPU(*args, fp64=True)
    v = np.random.default_rng(3).standard_normal(h.N)
    din = DeviceBuffer.from_array(v)
    dout = DeviceBuffer(v.nbytes)
    h.multiply_dev(din, dout)
    device_sync()
    w = np.empty_like(v)
    dout.download(w)
    w_ref = ref.apply(v)
    assert np.abs(w - w_ref).max() / np.abs(w_ref).max() < TOL
    w32 = h.multiply(v.astype(np.float32))
    assert w32.dtype == np.float32
    assert np.abs(w32 - ref.apply(v.astype(np.float32).astype(np.float64))).max() / np.abs(w_ref).max() < 1e-6


def test_batch_matches_single(models):
    from nushellx_amd.hv import Batch, DeviceBuffer, Hamiltonian, device_sync
    mf = models["sd44"]
    hams = [Hamiltonian(*mf.args(J2)) for J2 in mf.J2s]
    b = Batch(hams)
    assert b.N == sum(h.N for h in hams)
    v = np.random.default_rng(11).standard_normal(b.N)
    din = DeviceBuffer.from_array(v)
    dout = DeviceBuffer(v.nbytes)
    b.multiply_dev(din, dout)
    device_sync()
    w = np.empty_like(v)
    dout.download(w)
    ms = b.time(din, dout, warmup=3, iters=10)
    assert all(x > 0 for x in ms)
    # every tuned shape computes the same H*v
    cfg = b.autotune(iters=3)
    assert cfg["ms"] > 0
    for nt, mode, at in ((64, 0, 0), (128, 1, 0), (256, 0, 1), (256, 1, 1), (128, 0, 1)):
        b.set_config(dict(cfg, threads=nt, mode=mode, atomic=at))
        b.multiply_dev(din, dout)
        device_sync()
        w2 = np.empty_like(v)
        dout.download(w2)
        assert np.abs(w2 - w).max() <= 1e-13 * np.abs(w).max(), (nt, mode, at)
    for h, off in zip(hams, b.offsets):
        ws = h.multiply(v[off:off + h.N])
        assert np.abs(w[off:off + h.N] - ws).max() <= 1e-13 * max(np.abs(ws).max(), 1.0)


def test_trl_lowest_eigenvalues_match_dense(models):
    """Lowest 5 eigenvalues from the device TRL match dense eigvalsh of the
    brute-force H (refcpu unit-vector sweep, NuTRL's matrixs pattern)."""
    from nushellx_amd.hv import Hamiltonian
    for name, J2 in (("sd22", 0), ("sd22", 4), ("sd44", 0), ("sd44", 8), ("sd23", 3)):
        print(f"case {name} 2J={J2}", flush=True)
        args = models[name].args(J2)
        ref = RefCPU(*args, fp64=True)
        H = ref.dense()
        ev = np.linalg.eigvalsh(0.5 * (H + H.T))
        h = Hamiltonian(*args)
        nev = min(5, h.N)
        r = h.lanczos(nev=nev, max_basis=50, keep=10, tol=1e-8)
        assert r["converged"] == nev, r
        assert np.abs(r["evals"] - ev[:nev]).max() < 1e-6, (name, J2, r["evals"], ev[:nev])


def test_batch_lanczos_matches_dense_and_single(models):
    """Lockstep batch TRL: every instance's lowest 5 eigenvalues equal dense
    eigvalsh (small instances, some with fewer states than the Krylov basis,
    which exhaust their space) and the single-instance solver (larger ones)."""
    from nushellx_amd.hv import Batch, Hamiltonian
    mf = models["sd22"]
    hams = [Hamiltonian(*mf.args(J2)) for J2 in mf.J2s]
    nev = min(5, min(h.N for h in hams))
    r = Batch(hams).lanczos(nev=nev, max_basis=30, keep=10, tol=1e-9)
    assert (r["converged_at"] >= 0).all(), r
    for h, J2, ev in zip(hams, mf.J2s, r["evals"]):
        H = RefCPU(*mf.args(J2), fp64=True).dense()
        want = np.linalg.eigvalsh(0.5 * (H + H.T))[:nev]
        assert np.abs(ev - want).max() < 1e-7, (J2, ev, want)
    mf = models["sd44"]
    hams = [Hamiltonian(*mf.args(J2)) for J2 in mf.J2s]
    b = Batch(hams)
    rb = b.lanczos(nev=5, tol=1e-8, vectors=True)
    assert (rb["converged_at"] >= 0).all()
    for k, (h, ev) in enumerate(zip(hams, rb["evals"])):
        rs = h.lanczos(nev=5, tol=1e-8)
        assert np.abs(ev - rs["evals"]).max() < 1e-6
        o = b.offsets[k]
        for q in range(5):  # batch Ritz vectors are eigenvectors of their own instance
            z = rb["evecs"][q, o:o + h.N]
            assert abs(np.linalg.norm(z) - 1) < 1e-8
            assert np.linalg.norm(h.multiply(z) - ev[q] * z) < 1e-6


def test_lanczos_graph_replay_matches_stream_launches(models):
    """Blocks of Lanczos steps replayed as captured hipGraphs give the same
    eigenpairs as launching the same kernels on the stream."""
    from nushellx_amd.hv import Batch, Hamiltonian
    mf = models["sd44"]
    hams = [Hamiltonian(*mf.args(J2)) for J2 in mf.J2s]
    b = Batch(hams)
    rg = b.lanczos(nev=5, tol=1e-8, graph=True)
    rs = b.lanczos(nev=5, tol=1e-8, graph=False)
    rg2 = b.lanczos(nev=5, tol=1e-8, graph=True)
    assert rg["matvecs"] == rs["matvecs"] == rg2["matvecs"]
    assert np.abs(rg["evals"] - rs["evals"]).max() < 1e-9
    assert np.abs(rg2["evals"] - rs["evals"]).max() < 1e-9


def test_warm_start_converges_faster(models):
    from nushellx_amd.hv import Hamiltonian
    h = Hamiltonian(*models["sd44"].args(0))
    cold = h.lanczos(nev=1, tol=1e-8, vectors=True)
    warm = h.lanczos(nev=1, tol=1e-8, start=cold["evecs"][0])
    assert abs(warm["evals"][0] - cold["evals"][0]) < 1e-8
    assert warm["matvecs"] <= 8 < cold["matvecs"]


def test_solver_vectors_through_nuvec_files(models, tmp_path):
    """Device Ritz vectors written as .lnz/.abz: the states NuVec would form
    from them are eigenvectors of the oracle's H to REAL*4 precision."""
    from oracle.nuvec import nuvec_states
    from nushellx_amd.hv import Hamiltonian
    args = models["sd44"].args(0)
    h = Hamiltonian(*args)
    r = h.lanczos(nev=5, tol=1e-9, vectors=True)
    prefix = str(tmp_path / "sd440")
    h.write_vectors(prefix, r["evals"], r["evecs"], no_level=5)
    E, X = nuvec_states(prefix)
    ref = RefCPU(*args, fp64=True)
    assert np.abs(E - r["evals"]).max() < 1e-4   # REAL*4 energies
    for q in range(5):
        x = X[q] / np.linalg.norm(X[q])
        res = np.linalg.norm(ref.apply(x) - E[q] * x)
        assert res < 1e-4 * max(1.0, abs(E[q])), (q, res)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-4)])
def test_stacked_path_matches_refcpu(models, prec, tol):
    """Two-stage stacked path (X = V Bhat^T shared, then W += A sum r X) on a
    batch of every sd case (sd 6p6n has inner sizes above 24: ten K steps),
    for several chunk shapes, on the VALU and on the matrix cores, against the
    oracle."""
    from nushellx_amd.hv import Batch, DeviceBuffer, Hamiltonian, device_sync
    cases = [(n, J2) for n in ("sd22", "sd44", "sd66", "sd23") for J2 in models[n].J2s[:3]]
    hams = [Hamiltonian(*models[n].args(J2), precision=prec) for n, J2 in cases]
    refs = [RefCPU(*models[n].args(J2), fp64=True) for n, J2 in cases]
    rng = np.random.default_rng(21)
    for cfg in ({"stacked": 1, "threads": 256, "stk_cap": 2048, "stk_work": 512},
                {"stacked": 1, "threads": 128, "stk_cap": 4096, "stk_work": 256},
                {"stacked": 1, "threads": 256, "stk_cap": 1, "stk_work": 1},
                # on the matrix cores: a-terms per stage-2 workgroup, products per stage-1 wavefront
                {"stacked": 3, "threads": 256, "stk_cap": 8, "stk_work": 4},
                {"stacked": 3, "threads": 256, "stk_cap": 1, "stk_work": 1},
                {"stacked": 3, "threads": 256, "stk_cap": 64, "stk_work": 32},
                # both stages fused on the matrix cores (with the like-particle part): steps per
                # wavefront, output slots per task
                {"stacked": 4, "threads": 256, "stk_cap": 1, "stk_work": 2},
                {"stacked": 4, "threads": 256, "stk_cap": 2, "stk_work": 4},
                {"stacked": 4, "threads": 256, "stk_cap": 1000, "stk_work": 4}):
        print(f"case stacked {prec} {cfg}", flush=True)
        b = Batch(hams, precision=prec, config=dict(cfg, fat_cap=5120, max_phases=6, mode=0, atomic=1))
        if cfg["stacked"] == 4:
            assert b.path() == "fused-mfma +like"
        v = rng.standard_normal(b.N)
        din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
        b.multiply_dev(din, dout)
        device_sync()
        w = np.empty_like(v)
        dout.download(w)
        for k, r in enumerate(refs):
            o = b.offsets[k]
            ref = r.apply(v[o:o + r.N])
            err = np.abs(w[o:o + r.N] - ref).max() / np.abs(ref).max()
            assert err < tol, (cases[k], cfg, err)
        din.free()
        dout.free()
        b.close()


def test_trl_checkpoint_resume(models, tmp_path):
    """A solve cut short by max_iter leaves its last thick-restart state in the
    checkpoint file; a second solve resumes from it and ends with the eigenpairs
    and matvec count of an uninterrupted solve.  A checkpoint from another
    configuration is refused."""
    from nushellx_amd.hv import Batch, Hamiltonian, NsxError
    h = Hamiltonian(*models["sd44"].args(0))
    full = h.lanczos(nev=5, tol=1e-8)
    ck = str(tmp_path / "sd440.ckpt")
    cut = h.lanczos(nev=5, tol=1e-8, max_iter=60, checkpoint=ck)
    assert cut["converged"] == 0 and cut["matvecs"] < full["matvecs"]
    res = h.lanczos(nev=5, tol=1e-8, checkpoint=ck)
    assert res["converged"] == 5
    assert np.abs(res["evals"] - full["evals"]).max() < 1e-8
    assert abs(res["matvecs"] - full["matvecs"]) <= 8
    with pytest.raises(NsxError):
        h.lanczos(nev=5, keep=12, tol=1e-8, checkpoint=ck)
    # batches checkpoint every instance
    hams = [Hamiltonian(*models["sd44"].args(J2)) for J2 in (0, 4, 8)]
    b = Batch(hams)
    fb = b.lanczos(nev=5, tol=1e-8)
    ckb = str(tmp_path / "batch.ckpt")
    b.lanczos(nev=5, tol=1e-8, max_iter=60, checkpoint=ckb)
    rb = b.lanczos(nev=5, tol=1e-8, checkpoint=ckb)
    assert (rb["converged_at"] >= 0).all()
    assert np.abs(rb["evals"] - fb["evals"]).max() < 1e-8


def test_lanczos_basis_stays_orthonormal(models, tmp_path):
    """The classical Gram-Schmidt applied twice keeps the Krylov basis
    orthonormal: after the thick restarts of a batched solve cut short, the
    checkpointed restart basis of each instance (its
    kept Ritz vectors and the next Lanczos vector, read from the end of the
    checkpoint file) satisfies |Q^T Q - I| < 1e-12, and each kept vector is an
    approximate eigenvector of the oracle's H in the Rayleigh-quotient sense
    (Q^T H Q is diagonal up to the residual column)."""
    from nushellx_amd.hv import Batch, Hamiltonian
    cases = [("sd44", 0), ("sd44", 4), ("sd66", 0)]
    hams = [Hamiltonian(*models[n].args(J2)) for n, J2 in cases]
    b = Batch(hams)
    ck = str(tmp_path / "orth.ckpt")
    keep = 10
    r = b.lanczos(nev=5, max_basis=50, keep=keep, tol=1e-12, max_iter=130, checkpoint=ck)
    assert r["restarts"] >= 2
    raw = np.fromfile(ck, dtype=np.uint8)
    n, kk = b.N, keep
    Q = raw[raw.size - 8 * n * (kk + 1):].view(np.float64).reshape(kk + 1, n).T
    for k, (name, J2) in enumerate(cases):
        o, ni = b.offsets[k], hams[k].N
        Qi = Q[o:o + ni]
        G = Qi.T @ Qi
        err = np.abs(G - np.eye(kk + 1)).max()
        print(f"case orth {name} J2={J2}: |QtQ - I| = {err:.2e}", flush=True)
        assert err < 1e-12, (name, J2, err)
        ref = RefCPU(*models[name].args(J2), fp64=True)
        HQ = np.stack([ref.apply(Qi[:, c]) for c in range(kk)], axis=1)
        R = Qi[:, :kk].T @ HQ
        off = np.abs(R - np.diag(np.diag(R))).max()
        assert off < 1e-8 * np.abs(np.diag(R)).max(), (name, J2, off)


def test_standard_lanczos_matches_dense(models):
    """Standard Lanczos (no restarts, eigenvalue-stability test every 3 steps as
    Lnzs) reaches the dense eigenvalues; small spaces end exactly."""
    from nushellx_amd.hv import Hamiltonian
    for name, J2 in (("sd44", 0), ("sd22", 0)):
        args = models[name].args(J2)
        H = RefCPU(*args, fp64=True).dense()
        ev = np.linalg.eigvalsh(0.5 * (H + H.T))[:5]
        r = Hamiltonian(*args).lanczos(nev=5, max_basis=300, tol=1e-9, check_every=3, standard=True)
        print(name, r["matvecs"], r["evals"] - ev, flush=True)
        assert r["converged"] == 5
        assert np.abs(r["evals"] - ev).max() < 1e-7


def test_empty_and_ragged_instances(tmp_path):
    """A J with no basis states (sd 2p2n, 2J = 30): the handle has N = 0, its
    H*v is empty and a solve fails with an error; in a batch it sits between
    non-empty instances without moving them, and a batch solve, which would ask
    it for more states than it has, fails with an error."""
    from nushellx_amd.hv import Batch, DeviceBuffer, Hamiltonian, NsxError, device_sync
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import PNModel, write_model
    mf = write_model(PNModel(PNSpace(SD), 2, 2, random_interaction(SD, seed=3)), str(tmp_path), "sd22", [0, 4, 30])
    print("case empty instance sd22 2J=30", flush=True)
    h0 = Hamiltonian(*mf.args(30))
    assert h0.N == 0
    assert h0.multiply(np.zeros(0)).shape == (0,)
    with pytest.raises(NsxError, match="nothing to solve"):
        h0.lanczos(nev=1)
    hams = [Hamiltonian(*mf.args(0)), h0, Hamiltonian(*mf.args(4))]
    refs = [RefCPU(*mf.args(J2), fp64=True) for J2 in (0, 30, 4)]
    b = Batch(hams)
    assert b.offsets == [0, hams[0].N, hams[0].N]
    rng = np.random.default_rng(8)
    v = rng.standard_normal(b.N)
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    b.multiply_dev(din, dout)
    device_sync()
    w = np.empty_like(v)
    dout.download(w)
    for k, r in enumerate(refs):
        o = b.offsets[k]
        if r.N:
            ref = r.apply(v[o:o + r.N])
            assert np.abs(w[o:o + r.N] - ref).max() <= 1e-12 * np.abs(ref).max()
    with pytest.raises(NsxError, match="more eigenvalues"):
        b.lanczos(nev=1)
    din.free()
    dout.free()
    b.close()
