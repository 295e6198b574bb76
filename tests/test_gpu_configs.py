"""The batched ensemble configuration on a slice, against the oracle.

Uneven instances: sd 4p4n, 5p5n and 6p6n (block sides up to 33), J = 0, 2, 4,
for two interaction samples V_s = V_0 + N(0, (0.1 |V_0|)^2), as host-only
instances planned into one batch (18 instances, N from 1161 to 15089).  Every
instance's part of the batched H*v equals the FP64 oracle's H*v of its own
files, and the lockstep Lanczos of the batch gives every instance's lowest five
eigenvalues as the oracle's H has them (dense eigvalsh, or ARPACK driven by the
oracle's H*v): to 1e-6 MeV with FP64 contractions, to the 1e-4 MeV target with
the FP32 and BF16 hi/lo contractions."""
import tempfile

import numpy as np
import pytest

from oracle.refcpu import RefCPU
from tests.oracle_eigs import oracle_lowest

pytestmark = pytest.mark.gpu

J2S = [0, 4, 8]
NUCLEI = [(4, 4), (5, 5), (6, 6)]
SEEDS = [11, 12]


@pytest.fixture(scope="module")
def ens_slice():
    from nushellx_amd.ensemble import build_instances, write_ensemble
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    print("writing the ensemble slice", flush=True)
    base = random_interaction(SD, seed=2024)
    entries = write_ensemble(PNSpace(SD), NUCLEI, J2S, base, seeds=SEEDS,
                             workdir=tempfile.mkdtemp(prefix="nsx_slice_"))
    jobs = [(mf, J2) for (_, _, mf) in entries for J2 in J2S]
    hams = build_instances(entries, J2S, threads=8)
    print("oracle eigenvalues of the slice", flush=True)
    evs = [oracle_lowest(mf.args(J2)) for (mf, J2) in jobs]
    return {"jobs": jobs, "hams": hams, "evs": evs}


def test_slice_dims_and_hv_match_oracle(ens_slice):
    from nushellx_amd.hv import Batch, DeviceBuffer, device_sync
    jobs, hams = ens_slice["jobs"], ens_slice["hams"]
    dims = sorted({h.N for h in hams})
    assert dims[0] == 1161 and dims[-1] == 15089 and 2533 in dims  # 4p4n J=0, 6p6n J=4, 5p5n J=0
    b = Batch(hams)
    v = np.random.default_rng(3).standard_normal(b.N)
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    b.multiply_dev(din, dout)
    device_sync()
    w = np.empty_like(v)
    dout.download(w)
    for k, ((mf, J2), h) in enumerate(zip(jobs, hams)):
        o = b.offsets[k]
        ref = RefCPU(*mf.args(J2), fp64=True)
        assert ref.N == h.N
        w_ref = ref.apply(v[o:o + h.N])
        err = np.abs(w[o:o + h.N] - w_ref).max() / np.abs(w_ref).max()
        assert err < 1e-12, (k, J2, err)
        ref.close()
    b.close()


@pytest.mark.parametrize("prec,tol,ev_tol", [("fp64", 1e-7, 1e-6), ("fp32", 1e-5, 1e-4), ("bf16x3", 1e-5, 1e-4)])
def test_slice_eigenvalues_match_oracle(ens_slice, prec, tol, ev_tol):
    from nushellx_amd.hv import Batch
    print(f"case slice lanczos {prec}", flush=True)
    b = Batch(ens_slice["hams"], precision=prec)
    r = b.lanczos(nev=5, tol=tol)
    assert (r["converged_at"] >= 0).all()
    err = np.array([np.abs(r["evals"][k] - ev).max() for k, ev in enumerate(ens_slice["evs"])])
    print(f"{prec}: largest eigenvalue error per instance (MeV) {np.array2string(err, precision=2)}", flush=True)
    assert err.max() < ev_tol, (prec, err)
    # the samples differ: one nucleus and J has different spectra in the two samples
    n = len(NUCLEI) * len(J2S)
    assert np.abs(r["evals"][0] - r["evals"][n]).max() > 1e-3
    b.close()
