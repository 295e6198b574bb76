"""Linear ensembles on the GPU: one template per (nucleus, J) read from the base
interaction's files, every sample's operators made on the device from its
parameters.  The batch must equal the batch of per-sample file sets (the same
interactions written as NuShellX files, REAL*8) instance by instance: H*v in
every precision, the Lanczos eigenvalues, and again after new parameters; the
oracle checks the samples' files directly."""
import tempfile

import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu

NUCLEI = [(2, 2), (3, 3), (4, 4)]
J2S = [0, 4]
SEEDS = [101, 102, 103, 104, 105]


@pytest.fixture(scope="module")
def ens():
    from nushellx_amd.ensemble import build_instances, write_ensemble
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import SideCache
    space, base, cache = PNSpace(SD), random_interaction(SD, seed=77), SideCache()
    print("writing per-sample operator files", flush=True)
    entries = write_ensemble(space, NUCLEI, J2S, base, SEEDS, tempfile.mkdtemp(prefix="nsx_lin_files_"),
                             cache=cache, real_kind=8)
    hams = build_instances(entries, J2S, threads=8)
    return {"space": space, "base": base, "cache": cache, "entries": entries, "hams": hams}


def _hv(b, v):
    from nushellx_amd.hv import DeviceBuffer, device_sync
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    b.multiply_dev(din, dout)
    device_sync()
    w = np.empty_like(v)
    dout.download(w)
    din.free()
    dout.free()
    return w


def _linear(ens, precision="fp64", seeds=SEEDS, config=None):
    from nushellx_amd.ensemble import linear_ensemble
    print(f"case linear ensemble {precision} seeds={list(seeds)}", flush=True)
    return linear_ensemble(ens["space"], NUCLEI, J2S, ens["base"], seeds, tempfile.mkdtemp(prefix="nsx_lin_"),
                           precision=precision, config=config, cache=ens["cache"])


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_linear_ensemble_hv_equals_file_batch(ens):
    from nushellx_amd.hv import Batch
    ref = Batch(ens["hams"])
    lin, tmpl, _ = _linear(ens)
    assert lin.N == ref.N and lin.offsets == ref.offsets and lin.path() == "stacked"
    v = np.random.default_rng(3).standard_normal(ref.N)
    w_ref = _hv(ref, v)
    w = _hv(lin, v)
    for k, o in enumerate(ref.offsets):
        n = ens["hams"][k].N
        assert _rel(w[o:o + n], w_ref[o:o + n]) < 1e-12, k
    ref.close()
    lin.close()


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 2e-4)])
def test_linear_ensemble_reduced_precision(ens, prec, tol):
    from nushellx_amd.hv import Batch
    ref = Batch(ens["hams"])
    lin, _, _ = _linear(ens, prec)
    v = np.random.default_rng(5).standard_normal(ref.N)
    w_ref, w = _hv(ref, v), _hv(lin, v)
    for k, o in enumerate(ref.offsets):
        n = ens["hams"][k].N
        assert _rel(w[o:o + n], w_ref[o:o + n]) < tol, (prec, k)
    ref.close()
    lin.close()


def test_linear_ensemble_against_oracle(ens):
    """Two samples' instances against the oracle reading those samples' files."""
    lin, _, _ = _linear(ens)
    v = np.random.default_rng(9).standard_normal(lin.N)
    w = _hv(lin, v)
    jobs = [(mf, J2) for (_, _, mf) in ens["entries"] for J2 in J2S]
    nt = len(NUCLEI) * len(J2S)
    for k in list(range(nt)) + [3 * nt + 1, 4 * nt + nt - 1]:
        mf, J2 = jobs[k]
        r = RefCPU(*mf.args(J2), fp64=True)
        o = lin.offsets[k]
        assert _rel(w[o:o + r.N], r.apply(v[o:o + r.N])) < 1e-12, k
    lin.close()


def test_linear_ensemble_lanczos_and_update(ens):
    from nushellx_amd.hv import Batch
    ref = Batch(ens["hams"])
    r_ref = ref.lanczos(nev=5, tol=1e-8)
    ref.close()
    lin, tmpl, info = _linear(ens, seeds=[201, 202, 203, 204, 205])
    # new parameters for the same samples: the operators are rebuilt, the plan kept
    assert info["theta"].shape == (len(SEEDS), 66)
    lin.ens_update(np.stack([ens["base"].perturbed(s, 0.1).parameter_vector() for s in SEEDS]))
    r = lin.lanczos(nev=5, tol=1e-8)
    assert (r["converged_at"] >= 0).all() and (r_ref["converged_at"] >= 0).all()
    assert np.abs(r["evals"] - r_ref["evals"]).max() < 1e-7
    v = np.random.default_rng(13).standard_normal(lin.N)
    ref = Batch(ens["hams"])
    assert _rel(_hv(lin, v), _hv(ref, v)) < 1e-12
    ref.close()
    lin.close()


def test_linear_ensemble_with_tuned_config(ens):
    from nushellx_amd.hv import Batch
    small = Batch(ens["hams"][:len(NUCLEI) * len(J2S)])
    cfg = small.autotune(iters=3)
    small.close()
    cfg = dict(cfg, stacked=1)
    lin, _, _ = _linear(ens, config=cfg)
    ref = Batch(ens["hams"])
    v = np.random.default_rng(17).standard_normal(ref.N)
    assert _rel(_hv(lin, v), _hv(ref, v)) < 1e-12
    with pytest.raises(Exception):
        lin.autotune(iters=1)
    ref.close()
    lin.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-4)])
@pytest.mark.parametrize("stacked,path", [(3, "stacked-mfma"), (4, "fused-mfma")])
def test_linear_ensemble_on_matrix_core_paths(ens, prec, tol, stacked, path):
    """Every sample through the stacked stages on the matrix cores, two-stage
    and fused (grid y the sample; the fused path's untransposed Bhat copies made
    per sample on the device, its like-particle part in k_reduce), against the
    per-sample file batch in FP64, also after new parameters."""
    from nushellx_amd.hv import Batch
    ref = Batch(ens["hams"])
    cfg = {"fat_cap": 5120, "max_phases": 6, "threads": 256, "mode": 0, "atomic": 1, "stacked": stacked,
           "stk_cap": 8 if stacked == 3 else 1, "stk_work": 4}
    lin, _, _ = _linear(ens, prec, config=cfg)
    assert lin.path() == path
    v = np.random.default_rng(19).standard_normal(ref.N)
    w_ref, w = _hv(ref, v), _hv(lin, v)
    for k, o in enumerate(ref.offsets):
        n = ens["hams"][k].N
        assert _rel(w[o:o + n], w_ref[o:o + n]) < tol, (prec, k)
    ref.close()
    # the operators re-made for other parameters of the same samples
    lin.ens_update(np.stack([ens["base"].perturbed(s, 0.1).parameter_vector() for s in SEEDS]))
    ref = Batch(ens["hams"])
    w_ref, w = _hv(ref, v), _hv(lin, v)
    assert _rel(w, w_ref) < tol
    ref.close()
    lin.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 2e-4)])
def test_linear_ensemble_on_large_block_tiles(ens, monkeypatch, prec, tol):
    """Blocks with a side above 12 on the matrix-core tiles (every sample as the
    second grid dimension), against the per-sample file batch in FP64."""
    from nushellx_amd.hv import Batch
    ref = Batch(ens["hams"])
    v = np.random.default_rng(21).standard_normal(ref.N)
    w_ref = _hv(ref, v)
    ref.close()
    monkeypatch.setenv("NSX_BIG_MIN", "12")
    lin, _, _ = _linear(ens, prec)
    w = _hv(lin, v)
    for k, o in enumerate(lin.offsets):
        n = ens["hams"][k].N
        assert _rel(w[o:o + n], w_ref[o:o + n]) < tol, (prec, k)
    lin.close()


def test_linear_ensemble_pf44():
    """pf 4p4n J=0 (blocks up to 99 on the large-block tiles): three samples
    against their own file sets."""
    from nushellx_amd.ensemble import build_instances, linear_ensemble, write_ensemble
    from nushellx_amd.hv import Batch
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import PF, PNSpace
    from nushellx_amd.synth.writer import SideCache
    space, cache = PNSpace(PF), SideCache()
    base = random_interaction(PF, seed=2024, spe_range=(-10.0, -4.0))
    seeds = [301, 302, 303]
    print("case linear ensemble pf44: per-sample files", flush=True)
    entries = write_ensemble(space, [(4, 4)], [0], base, seeds, tempfile.mkdtemp(prefix="nsx_pf_files_"),
                             cache=cache, real_kind=8)
    ref = Batch(build_instances(entries, [0], threads=4))
    print("case linear ensemble pf44: linear", flush=True)
    lin, _, _ = linear_ensemble(space, [(4, 4)], [0], base, seeds, tempfile.mkdtemp(prefix="nsx_pf_lin_"),
                                cache=cache)
    assert lin.N == ref.N
    v = np.random.default_rng(31).standard_normal(ref.N)
    assert _rel(_hv(lin, v), _hv(ref, v)) < 1e-12
    r = lin.lanczos(nev=3, tol=1e-6)
    r_ref = ref.lanczos(nev=3, tol=1e-6)
    assert np.abs(r["evals"] - r_ref["evals"]).max() < 1e-6
    ref.close()
    lin.close()
