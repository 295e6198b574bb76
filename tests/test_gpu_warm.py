"""Warm-start network on the GPU: gradients and outputs of the HIP kernels
against the CPU oracle (oracle/mlp.py, float64 with the same BF16-rounded
weights), training, and the gain of the trained network over a random start."""
import os
import tempfile

import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


# the last two have an output layer long enough that the input-gradient
# product of the output layer is split over K (partials summed in order)
@pytest.mark.parametrize("shape", [(6, 32, 2, 10, 7), (66, 512, 3, 300, 40), (17, 64, 1, 1000, 130),
                                   (20, 64, 2, 5000, 24), (33, 128, 1, 70000, 16)])
def test_mlp_gradients_and_outputs_match_oracle(shape):
    from nushellx_amd.warm import MLP
    n_in, hid, layers, n_out, rows = shape
    print(f"case mlp {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    net = MLP(n_in, n_out, hid, layers, seed=7)
    p = net.get_params()
    p[:] += rng.normal(size=p.size).astype(np.float32) * 0.02  # non-zero biases too
    net.set_params(p)
    X = rng.normal(size=(rows, n_in)).astype(np.float32)
    Y = rng.normal(size=(rows, n_out)).astype(np.float32)
    loss, g = net.grad(X, Y)
    lo, go = omlp.loss_grad(p, net.widths, X, Y, quantize=True)
    assert abs(loss - lo) <= 1e-5 * lo
    for w_off, b_off, no, ni in omlp.layout(net.widths):  # per layer, FP32 vs float64
        for a, b in ((w_off, w_off + no * ni), (b_off, b_off + no)):
            ref = go[a:b]
            assert np.abs(g[a:b] - ref).max() <= 2e-5 * np.abs(ref).max() + 1e-30, (a, b)
    yp = net.predict(X)
    yo, _ = omlp.forward(p, net.widths, X, quantize=True)
    assert np.abs(yp - yo).max() <= 1e-5 * np.abs(yo).max()
    net.close()


def test_mlp_training_fits_a_smooth_map():
    from nushellx_amd.warm import MLP
    rng = np.random.default_rng(1)
    X = rng.normal(size=(256, 8)).astype(np.float32)
    A = rng.normal(size=(8, 40)) / np.sqrt(8)
    Y = (np.tanh(X @ A) + 0.3 * (X @ A)).astype(np.float32)
    net = MLP(8, 40, 128, 2, seed=2)
    l0 = net.grad(X, Y)[0]
    losses = net.train(X, Y, epochs=150, batch=32, lr=2e-3, seed=3)
    assert losses[-1] < 0.02 * l0, (l0, losses[::30])
    assert abs(net.grad(X, Y)[0] - losses[-1]) < 0.5 * losses[-1] + 1e-3


def test_warm_start_gain_over_random_start():
    """sd 4p4n, J = 0, 2, 4: network trained on the lowest-8 Lanczos
    eigenvectors of 256 perturbed interactions (V_s = V_0 + N(0, (0.1 |V_0|)^2)).
    On held-out interactions the predicted start vector lies closer to the
    wanted subspace than the training mean does, the warm-started batch needs
    fewer H*v than the random start in total, and the eigenvalues agree."""
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import PNModel, SideCache, write_model
    from nushellx_amd.warm import WarmStart, build_training_set, solve_batch
    J2s = [0, 4, 8]
    base = random_interaction(SD, seed=2024)
    work = tempfile.mkdtemp(prefix="nsx_warm_")
    cache = SideCache()
    F, T, ref, _ = build_training_set(PNSpace(SD), 4, 4, J2s, base, range(1000, 1256), work, cache=cache,
                                      log=lambda m: print(m, flush=True))
    assert len(F) >= 240
    ws = WarmStart(F.shape[1], T.shape[1], seed=0)
    losses = ws.fit(F, T, epochs=200, batch=16, lr=1e-3)
    print("fit loss", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.2 * losses[0]
    n_cold = n_warm = 0
    leak_mlp, leak_mean = [], []
    for s in (5001, 5002, 5003):
        inter = base.perturbed(s)
        mf = write_model(PNModel(PNSpace(SD), 4, 4, inter), work, "ws", J2s, cache=cache, densities=False,
                         xtra="y")
        exact, b, hams = solve_batch(mf, J2s, nev=8, tol=1e-8, vectors=True)
        cold = b.lanczos(nev=5, tol=1e-4, check_every=4)
        v, fell = ws.start_vector(b, inter.parameter_vector()[None, :])
        warm = b.lanczos(nev=5, tol=1e-4, start=v, check_every=4)
        for o, n in zip(exact["offsets"], exact["dims"]):
            P = exact["evecs"][:, o:o + n]
            for x, acc in ((v[o:o + n], leak_mlp), (ws.t_mean[o:o + n], leak_mean)):
                x = x / np.linalg.norm(x)
                acc.append(np.sqrt(max(0.0, 1.0 - np.sum((P @ x) ** 2))))
        b.close()
        print(f"held-out {s}: random {cold['matvecs']} H*v, warm {warm['matvecs']} H*v, fallback {fell}",
              flush=True)
        assert np.abs(warm["evals"] - exact["evals"][:, :5]).max() < 1e-4
        assert np.abs(cold["evals"] - exact["evals"][:, :5]).max() < 1e-4
        n_cold += cold["matvecs"]
        n_warm += warm["matvecs"]
    print("leakage mlp", np.round(leak_mlp, 3), "mean", np.round(leak_mean, 3), flush=True)
    assert np.mean(leak_mlp) < np.mean(leak_mean)
    assert n_warm < n_cold
