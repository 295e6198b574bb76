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
                   prefix: str = "e", log=None, real_kind: int = 4):
    """Operator files for every (sample, nucleus).  Returns a list of
    (sample seed, (np, nn), ModelFiles) in sample-major order."""
    from .synth.writer import PNModel, SideCache, write_model
    cache = cache or SideCache()
    out = []
    base_dirs = {}
    for (np_, nn) in nuclei:
        name = f"{prefix}{np_}{nn}"
        d = os.path.join(workdir, name)
        write_model(PNModel(space, np_, nn, base_inter), d, name, list(J2s), cache=cache, real_kind=real_kind)
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
                             densities=False, real_kind=real_kind)
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


# ---------------------------------------------------------------- linear ensembles
def template_basis(ham, space, model, cache, oph_basis=None) -> np.ndarray:
    """P x K basis of a template handle: its values (``Hamiltonian.ens_layout``)
    for every unit parameter vector of ``model.inter``, from the generator's
    linear forms (no files)."""
    from .synth.interaction import multipole_pn_basis
    no, hl = ham.ens_layout()
    keys, _, M = oph_basis if oph_basis is not None else multipole_pn_basis(space, model.inter)
    if no and M.shape[1] != no:
        raise ValueError(f"template has {no} .oph records, the interaction {M.shape[1]}")
    ra, rb = model.roles()
    parts = [M if no else np.zeros((M.shape[0], 0))]
    dense = {}
    for side, J2, d in hl:
        kind = ra if side == 0 else rb
        n = model.nprot if kind == "p" else model.nneut
        if (kind, n) not in dense:
            dense[(kind, n)] = cache.h_dense_basis(space, kind, n, model.inter)
        B = dense[(kind, n)][J2]
        if B.shape[1] != d * d:
            raise ValueError(f"side {side} 2J={J2}: template d={d}, generator d^2={B.shape[1]}")
        parts.append(B)
    return np.ascontiguousarray(np.concatenate(parts, axis=1))


def ensemble_templates(space, nuclei: Sequence[Tuple[int, int]], J2s: Sequence[int], base_inter, workdir: str,
                       cache=None, prefix: str = "e"):
    """Templates of a linear ensemble, one per (nucleus, J) in nucleus-major
    order, read from files of the base interaction (written here), and their
    bases.  Returns (templates, bases, timings: template_files_s for the
    generator writing the files, densities included; templates_s for reading
    the templates and forming the bases)."""
    from .hv import Hamiltonian
    from .synth.interaction import multipole_pn_basis
    from .synth.writer import PNModel, SideCache, write_model
    cache = cache or SideCache()
    t0 = time.perf_counter()
    files = []
    for (np_, nn) in nuclei:  # the base interaction's files: densities, H(J), multipoles
        name = f"{prefix}{np_}{nn}"
        model = PNModel(space, np_, nn, base_inter)
        files.append((model, write_model(model, os.path.join(workdir, name), name, list(J2s), cache=cache,
                                         real_kind=8)))
    t_files = time.perf_counter() - t0
    t0 = time.perf_counter()
    oph = multipole_pn_basis(space, base_inter)
    templates, bases = [], []
    for model, mf in files:
        B = None
        for J2 in J2s:
            h = Hamiltonian(*mf.args(J2), device=False, template=True)
            if B is None:  # every J of a nucleus reads the same side and interaction files
                B = template_basis(h, space, model, cache, oph_basis=oph)
            templates.append(h)
            bases.append(B)
    return templates, bases, {"template_files_s": t_files, "templates_s": time.perf_counter() - t0}


def linear_ensemble(space, nuclei: Sequence[Tuple[int, int]], J2s: Sequence[int], base_inter,
                    seeds: Sequence[int], workdir: str, rel_sigma: float = 0.1, precision: str = "fp64",
                    config: Optional[dict] = None, cache=None, prefix: str = "e", log=None, templates=None):
    """The ensemble of ``write_ensemble`` (samples V_0 + N(0, (rel_sigma |V_0|)^2)
    of ``base_inter`` for every nucleus and J, instances sample-major, then
    nucleus, then J) without a file set per sample: one template per (nucleus,
    J) read from files of the base interaction, the samples' parameters and the
    templates' bases handed to ``Batch.linear_ensemble``.  ``templates``: the
    result of ``ensemble_templates`` to reuse.  Returns (batch, templates,
    timings: those of ``ensemble_templates``, theta_s, batch_s for the plan,
    upload and every sample's operators made on the device)."""
    from .hv import Batch
    if templates is None:
        templates = ensemble_templates(space, nuclei, J2s, base_inter, workdir, cache=cache, prefix=prefix)
    tmpl, bases, info = templates
    info = dict(info)
    t0 = time.perf_counter()
    theta = np.stack([base_inter.perturbed(int(s), rel_sigma).parameter_vector() for s in seeds])
    info["theta_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = Batch.linear_ensemble(tmpl, theta, bases, precision=precision, config=config)
    info["batch_s"] = time.perf_counter() - t0
    info["theta"] = theta
    if log:
        log(f"linear ensemble: {len(tmpl)} templates x {len(seeds)} samples, template files "
            f"{info['template_files_s']:.2f} s, templates {info['templates_s']:.2f} s, batch {info['batch_s']:.2f} s")
    return b, tmpl, info
