"""Learned warm start on pf 4p4n, J = 0 (N = 41,355; 199 interaction parameters).

The 3 x 512 network (BF16 weights, FP32 accumulation, trained by the HIP
kernels) is fitted to Lanczos ground-state targets (the lowest eight states of
each perturbed interaction V_s = V_0 + N(0, (0.1 |V_0|)^2), solved 32 at a time
as one batch).  On held-out interactions the warm-started solve needs fewer H*v
than the random start, a prediction that is no better than a random vector is
replaced by one (the fallback), and the warm-started eigenvalues equal the
lowest five of the oracle's H (ARPACK on the oracle's H*v)."""
import tempfile

import numpy as np
import pytest

from tests.oracle_eigs import oracle_lowest

pytestmark = pytest.mark.gpu

TRAIN = 64


@pytest.fixture(scope="module")
def pf_ws():
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import PF, PNSpace
    from nushellx_amd.warm import WarmStart, build_training_set_batched
    base = random_interaction(PF, seed=2024, spe_range=(-10.0, -4.0))
    work = tempfile.mkdtemp(prefix="nsx_pfws_")
    print(f"pf 4p4n J=0 training set: {TRAIN} perturbed interactions", flush=True)
    F, T, ref, _ = build_training_set_batched(PNSpace(PF), 4, 4, [0], base, range(3000, 3000 + TRAIN), work,
                                              nev=8, tol=1e-5, chunk=32, threads=8,
                                              log=lambda m: print(m, flush=True))
    assert len(F) >= TRAIN - 4 and F.shape[1] == 199 and T.shape[1] == 41355
    ws = WarmStart(F.shape[1], T.shape[1], hidden=512, layers=3, seed=0)
    losses = ws.fit(F, T, epochs=150, batch=16, lr=1e-3)
    print("fit loss", losses[0], losses[-1], flush=True)
    return {"base": base, "ws": ws, "work": work, "losses": losses}


def test_pf44_warm_start_gain_fallback_and_oracle(pf_ws):
    from nushellx_amd.synth.mspace import PF, PNSpace
    from nushellx_amd.synth.writer import PNModel, SideCache, write_model
    from nushellx_amd.hv import Batch, Hamiltonian
    base, ws, work = pf_ws["base"], pf_ws["ws"], pf_ws["work"]
    assert pf_ws["losses"][-1] < 0.5 * pf_ws["losses"][0]
    cache = SideCache()
    n_cold = n_warm = 0
    for k, s in enumerate((7001, 7002, 7003)):
        inter = base.perturbed(s)
        mf = write_model(PNModel(PNSpace(PF), 4, 4, inter), f"{work}/h{k}", "pfh", [0], cache=cache)
        hams = [Hamiltonian(*mf.args(0))]
        b = Batch(hams)
        cold = b.lanczos(nev=5, tol=1e-5, check_every=4)
        v, fell = ws.start_vector(b, inter.parameter_vector()[None, :])
        assert fell == []
        warm = b.lanczos(nev=5, tol=1e-5, check_every=4, start=v)
        print(f"held-out {s}: converged at {cold['converged_at'][0]} (random) vs {warm['converged_at'][0]} (warm) H*v",
              flush=True)
        assert np.abs(warm["evals"] - cold["evals"]).max() < 1e-5
        n_cold += int(cold["converged_at"][0])
        n_warm += int(warm["converged_at"][0])
        if k == 0:
            # eigenvalues of the warm-started solve against the oracle's H
            tight = b.lanczos(nev=5, tol=1e-7, check_every=4, start=v)
            ev = oracle_lowest(mf.args(0))
            print("warm-started", tight["evals"][0], "oracle", ev, flush=True)
            assert np.abs(tight["evals"][0] - ev).max() < 1e-6
            # fallback: a prediction with no usable direction is replaced by the random start
            ws_bad = ws.predict
            ws.predict = lambda f: np.full((1, b.N), np.nan)
            try:
                vb, fell_b = ws.start_vector(b, inter.parameter_vector()[None, :])
            finally:
                ws.predict = ws_bad
            assert fell_b == [0] and np.isfinite(vb).all()
            fb = b.lanczos(nev=5, tol=1e-5, check_every=4, start=vb)
            assert np.abs(fb["evals"] - cold["evals"]).max() < 1e-5
        b.close()
        for h in hams:
            h.close()
    print(f"H*v to convergence over the held-out set: random {n_cold}, warm {n_warm}", flush=True)
    assert n_warm < n_cold
