"""Ensembles: many interaction samples x nuclei x J values solved as one batch.

Every sample V_s = V_0 + eps (eps_i ~ N(0, (rel_sigma |V_0,i|)^2)) gets its own
NuShellX operator files, written in a directory of its own.  The density files
(.tri/.trd) do not depend on the interaction, so each sample directory links to
the nucleus' single copy.  The instances are built host-only (no device memory
each) in a thread pool and planned once as one batch.  The batch then runs
every instance in lockstep through the thick-restart Lanczos: one batched H*v
per step, one captured graph per block of steps.  Instances that share density
blocks (all samples of a nucleus) store them once in HBM.
"""
from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor
from typing import List, Optional, Sequence, Tuple

import numpy as np


def write_ensemble(space, nuclei: Sequence[Tuple[int, int]], J2s: Sequence[int], base_inter,
                   seeds: Sequence[int], workdir: str, rel_sigma: float = 0.1, cache=None,
                   prefix: str = "e", log=None):
    """Operator files for every (sample, nucleus).  Returns a list of
    (sample seed, (np, nn), ModelFiles) in sample-major order."""
    from .synth.writer import PNModel, SideCache, write_model
    cache = cache or SideCache()
    out = []
    base_dirs = {}
    for (np_, nn) in nuclei:
        name = f"{prefix}{np_}{nn}"
        d = os.path.join(workdir, name)
        write_model(PNModel(space, np_, nn, base_inter), d, name, list(J2s), cache=cache)
        base_dirs[(np_, nn)] = (d, name)
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        inter = base_inter.perturbed(int(s), rel_sigma)
        for (np_, nn) in nuclei:
            d, name = base_dirs[(np_, nn)]
            sd = os.path.join(d, f"s{k:05d}")
            os.makedirs(sd, exist_ok=True)
            for role in ("a", "b"):
                for ext in (".tri", ".trd"):
                    link = os.path.join(sd, name + role + ext)
                    if not os.path.lexists(link):
                        os.symlink(os.path.join("..", name + role + ext), link)
            mf = write_model(PNModel(space, np_, nn, inter), sd, name, list(J2s), cache=cache,
                             densities=False)
            out.append((int(s), (np_, nn), mf))
        if log and (k + 1) % 64 == 0:
            log(f"ensemble files: {k + 1}/{len(seeds)} samples ({time.perf_counter() - t0:.0f} s)")
    return out


def build_instances(entries, J2s: Sequence[int], threads: int = 16, device: bool = False):
    """Host-only Hamiltonians for every (entry, J) in entry-major order."""
    from .hv import Hamiltonian
    jobs = [(mf, J2) for (_, _, mf) in entries for J2 in J2s]
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda a: Hamiltonian(*a[0].args(a[1]), device=device), jobs))


def solve_ensemble(hams, nev: int = 5, tol: float = 1e-4, precision: str = "fp64", tune_on: Optional[list] = None,
                   log=None, **kw):
    """One batch over all instances; returns the Lanczos result plus timings
    (tuning, plan/upload, solve) and eigenpairs per second of the solve.
    ``tune_on``: a few representative instances (one sample of every nucleus
    and J) whose small batch the autotuner times; the big batch is then planned
    once with the winning shape."""
    from .hv import Batch
    tuned, t_tune = None, 0.0
    if tune_on:
        t0 = time.perf_counter()
        small = Batch(tune_on, precision=precision)
        tuned = small.autotune(iters=5)
        small.close()
        t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = Batch(hams, precision=precision, config=tuned)
    t_plan = time.perf_counter() - t0
    if log:
        log(f"ensemble batch: {len(hams)} instances, N_total {b.N}, tune {t_tune:.1f} s, plan {t_plan:.1f} s")
    t0 = time.perf_counter()
    r = b.lanczos(nev=nev, tol=tol, **kw)
    t_solve = time.perf_counter() - t0
    r.update({"batch": b, "t_plan": t_plan, "t_tune": t_tune, "t_solve": t_solve, "tuned": tuned,
              "eigenpairs_per_s": len(hams) * nev / t_solve, "N_total": b.N})
    return r
