"""Learned warm start for the Lanczos solver.

A multilayer perceptron (trained and evaluated by the HIP kernels of
``csrc/nsx_mlp.hip``) maps the interaction parameters of a model space (SPEs
and JT two-body matrix elements) to a Lanczos start vector for every instance
(J value) of a batch, in the partition-block basis order of the H*v engine.

Training targets are produced by the solver itself: for each perturbed
interaction, the lowest ``nev`` Lanczos eigenvectors of every instance are
sign-aligned to those of a reference interaction and summed (normalised per
instance).  A start vector in the span of the wanted states makes the Krylov
space capture all of them within a few steps.

The random-start thick-restart Lanczos stays the fallback and the arbiter: a
predicted start vector whose Rayleigh-quotient residual is not below that of a
random vector, per instance, is replaced by a random one; the solver's own
residual test decides convergence either way.  The reference has no such
component (its start is 0.5 - U on the parity mask, src/NuTRL/NuCoreX.f90:299-312).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from ._lib import check, lib


class MLP:
    """n_in -> n_hidden (SiLU) x n_layers -> n_out; BF16 weights, FP32 accumulation."""

    def __init__(self, n_in: int, n_out: int, n_hidden: int = 512, n_layers: int = 3, seed: int = 0):
        h = C.c_void_p()
        check(lib().nsx_mlp_create(n_in, n_hidden, n_layers, n_out, seed, C.byref(h)), "nsx_mlp_create")
        self._h = h
        self.widths = [n_in] + [n_hidden] * n_layers + [n_out]
        n = C.c_int64()
        check(lib().nsx_mlp_num_params(h, C.byref(n)), "nsx_mlp_num_params")
        self.n_params = int(n.value)

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        check(lib().nsx_mlp_get_params(self._h, p.ctypes.data), "nsx_mlp_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        check(lib().nsx_mlp_set_params(self._h, p.ctypes.data), "nsx_mlp_set_params")

    @staticmethod
    def _rows(X, width):
        X = np.ascontiguousarray(X, dtype=np.float32)
        assert X.ndim == 2 and X.shape[1] == width, (X.shape, width)
        return X

    def grad(self, X, Y):
        X, Y = self._rows(X, self.widths[0]), self._rows(Y, self.widths[-1])
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        check(lib().nsx_mlp_grad(self._h, X.ctypes.data, Y.ctypes.data, X.shape[0], g.ctypes.data,
                                 C.byref(loss)), "nsx_mlp_grad")
        return float(loss.value), g

    def train(self, X, Y, epochs: int, batch: int = 32, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        X, Y = self._rows(X, self.widths[0]), self._rows(Y, self.widths[-1])
        losses = np.zeros(epochs, dtype=np.float32)
        check(lib().nsx_mlp_train(self._h, X.ctypes.data, Y.ctypes.data, X.shape[0], epochs, batch,
                                  C.c_float(lr), seed, losses.ctypes.data), "nsx_mlp_train")
        return losses

    def predict(self, X) -> np.ndarray:
        X = self._rows(X, self.widths[0])
        Y = np.zeros((X.shape[0], self.widths[-1]), dtype=np.float32)
        check(lib().nsx_mlp_predict(self._h, X.ctypes.data, X.shape[0], Y.ctypes.data), "nsx_mlp_predict")
        return Y

    def close(self):
        if getattr(self, "_h", None):
            lib().nsx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def align_and_sum(vecs: np.ndarray, ref: np.ndarray, offsets: Sequence[int], dims: Sequence[int]) -> np.ndarray:
    """Target vector of one sample: per instance, each wanted eigenvector is
    given the sign of its overlap with the best-matching reference eigenvector,
    and the signed vectors are summed and normalised.  vecs, ref: nev x N_total."""
    out = np.zeros(vecs.shape[1])
    for o, n in zip(offsets, dims):
        R = ref[:, o:o + n]
        acc = np.zeros(n)
        for q in range(vecs.shape[0]):
            v = vecs[q, o:o + n]
            ov = R @ v
            p = int(np.argmax(np.abs(ov)))
            acc += v if ov[p] >= 0 else -v
        nrm = np.linalg.norm(acc)
        out[o:o + n] = acc / nrm if nrm > 0 else acc
    return out


class WarmStart:
    """Standardised features -> MLP -> de-standardised start vectors."""

    def __init__(self, n_features: int, n_total: int, hidden: int = 512, layers: int = 3, seed: int = 0):
        self.net = MLP(n_features, n_total, hidden, layers, seed)
        self.n_total = n_total

    def fit(self, features: np.ndarray, targets: np.ndarray, epochs: int = 200, batch: int = 32,
            lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        F = np.asarray(features, dtype=np.float64)
        T = np.asarray(targets, dtype=np.float64)
        self.f_mean = F.mean(axis=0)
        sd = F.std(axis=0)
        self.f_std = np.where(sd > 1e-12, sd, 1.0)
        self.t_mean = T.mean(axis=0)
        self.t_scale = float(np.sqrt(np.mean((T - self.t_mean) ** 2))) or 1.0
        X = (F - self.f_mean) / self.f_std
        Y = (T - self.t_mean) / self.t_scale
        return self.net.train(X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed)

    def predict(self, features: np.ndarray) -> np.ndarray:
        F = np.atleast_2d(np.asarray(features, dtype=np.float64))
        X = (F - self.f_mean) / self.f_std
        return self.net.predict(X).astype(np.float64) * self.t_scale + self.t_mean

    def start_vector(self, batch, features: np.ndarray, seed: int = 99) -> tuple:
        """Predicted start vector for a Batch, with the random fallback applied
        per instance.  ``features``: one row per sample; the batch holds the
        samples' instances sample-major, in the order the network was trained
        on.  Returns (vector, list of instances that fell back)."""
        from .hv import DeviceBuffer, device_sync
        v = self.predict(features).reshape(-1)
        if v.size != batch.N:
            raise ValueError(f"prediction has {v.size} amplitudes, the batch {batch.N}")
        rng = np.random.default_rng(seed)
        r = np.zeros_like(v)
        fell = []
        for o, h in zip(batch.offsets, batch._hams):
            r[o:o + h.N] = (0.5 - rng.random(h.N)) * h.lcoef()
        # one batched H*v per candidate: Rayleigh-quotient residuals
        din = DeviceBuffer.from_array(v)
        dout = DeviceBuffer(v.nbytes)
        hv, hr = np.empty_like(v), np.empty_like(r)
        batch.multiply_dev(din, dout)
        device_sync()
        dout.download(hv)
        din.upload(r)
        batch.multiply_dev(din, dout)
        device_sync()
        dout.download(hr)
        din.free()
        dout.free()
        out = v.copy()
        for k, (o, h) in enumerate(zip(batch.offsets, batch._hams)):
            sl = slice(o, o + h.N)

            def rq_res(x, hx):
                n2 = x @ x
                if not n2 > 0 or not np.isfinite(n2):
                    return np.inf
                th = (x @ hx) / n2
                return np.linalg.norm(hx - th * x) / np.sqrt(n2)
            if not rq_res(v[sl], hv[sl]) < rq_res(r[sl], hr[sl]):
                out[sl] = r[sl]
                fell.append(k)
        return out, fell


def solve_batch(mf, J2s, nev: int = 5, tol: float = 1e-6, precision: str = "fp64", start=None,
                vectors: bool = True, autotune: bool = False, **kw):
    """Lanczos on every J instance of a written model in one lockstep batch."""
    from .hv import Batch, Hamiltonian
    hams = [Hamiltonian(*mf.args(J2), precision=precision) for J2 in J2s]
    b = Batch(hams, precision=precision)
    if autotune:
        b.autotune(iters=5)
    r = b.lanczos(nev=nev, tol=tol, vectors=vectors, start=start, **kw)
    r["offsets"] = list(b.offsets)
    r["dims"] = [h.N for h in hams]
    return r, b, hams


def _linear_training_set(space, nprot, nneut, J2s, base_inter, seeds, workdir, name, nev, tol, rel_sigma, chunk,
                         cache, log):
    """Training targets from linear ensembles: the base interaction's files are
    written once (densities included) and read as one template per J; the
    samples are solved ``chunk`` at a time as one lockstep batch whose operators
    are made on the device from their parameters (no files per sample)."""
    from .ensemble import template_basis
    from .hv import Batch, Hamiltonian
    from .synth.writer import PNModel, SideCache, write_model
    cache = cache or SideCache()
    model = PNModel(space, nprot, nneut, base_inter)
    mf = write_model(model, workdir, name, list(J2s), cache=cache, real_kind=8)
    tmpl = [Hamiltonian(*mf.args(J2), device=False, template=True) for J2 in J2s]
    basis = template_basis(tmpl[0], space, model, cache)  # every J reads the same side and interaction files
    bases = [basis] * len(tmpl)
    refb = Batch(tmpl)
    ref = refb.lanczos(nev=nev, tol=tol, vectors=True)
    ref["offsets"] = list(refb.offsets)
    ref["dims"] = dims = [h.N for h in tmpl]
    refb.close()
    nJ = len(J2s)
    feats, targs = [], []
    seeds = list(seeds)
    for c0 in range(0, len(seeds), chunk):
        part = seeds[c0:c0 + chunk]
        theta = np.stack([base_inter.perturbed(int(s), rel_sigma).parameter_vector() for s in part])
        b = Batch.linear_ensemble(tmpl, theta, bases)
        r = b.lanczos(nev=nev, tol=tol, vectors=True)
        for k in range(len(part)):
            if (r["converged_at"][k * nJ:(k + 1) * nJ] < 0).any():
                continue  # unconverged samples are not used as targets
            o0 = b.offsets[k * nJ]
            vec = r["evecs"][:, o0:o0 + sum(dims)]
            offs = [b.offsets[k * nJ + j] - o0 for j in range(nJ)]
            feats.append(theta[k])
            targs.append(align_and_sum(vec, ref["evecs"], offs, dims))
        b.close()
        if log:
            log(f"training set: {min(c0 + chunk, len(seeds))}/{len(seeds)} samples solved")
    for h in tmpl:
        h.close()
    return np.array(feats), np.array(targs), ref["evecs"], ref


def build_training_set(space, nprot: int, nneut: int, J2s: Sequence[int], base_inter, seeds: Sequence[int],
                       workdir: str, name: str = "ws", nev: int = 8, tol: float = 1e-6, rel_sigma: float = 0.1,
                       cache=None, log=None, chunk: int = 256):
    """Features and targets from Lanczos solves of perturbed interactions.

    Targets sum the lowest ``nev`` eigenvectors; taking a few more states than
    are wanted keeps a level crossing just above the wanted ones from flipping
    the target between neighbouring samples.  The operator files of the
    reference interaction are written once (in ``workdir`` under ``name``,
    densities included, so that further interactions can be written beside
    them without densities); the samples are solved ``chunk`` at a time as a
    linear ensemble.  Returns (features, targets, reference eigenvectors,
    reference solve)."""
    return _linear_training_set(space, nprot, nneut, J2s, base_inter, seeds, workdir, name, nev, tol, rel_sigma,
                                chunk, cache, log)


def build_training_set_batched(space, nprot: int, nneut: int, J2s: Sequence[int], base_inter,
                               seeds: Sequence[int], workdir: str, nev: int = 8, tol: float = 1e-6,
                               rel_sigma: float = 0.1, chunk: int = 64, threads: int = 16, log=None):
    """As ``build_training_set`` with a smaller default chunk (large model
    spaces); ``threads`` is kept for callers and unused."""
    return _linear_training_set(space, nprot, nneut, J2s, base_inter, seeds, workdir, "ws", nev, tol, rel_sigma,
                                chunk, None, log)
