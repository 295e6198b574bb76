#!/usr/bin/env python3
"""Benchmark: batched pn J-scheme H*v (NuTRL's multiplyx/a/b) on one MI355X.

Workload: sd shell, 4 valence protons + 4 valence neutrons, every J = 0..8
(nine H*v instances, N = 27,837 amplitudes in total), seeded isospin-conserving
interaction (63 JT TBMEs ~ N(0, 1.5 MeV), SPEs ~ U[-4, 0] MeV), operator files
generated in NuShellX format before timing.  One step = one H*v of every
instance (one launch pair).  ``value`` = H*v applications per second with the
vectors resident in HBM.  The CPU baseline is the C/OpenMP restatement of the
reference loop nest in its REAL*4 build, timed on this host's cores on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
SEED = 2024
HELD_OUT_SEED = 777


def make_workload(outdir: str):
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import PNModel, write_model
    inter = random_interaction(SD, seed=SEED)
    return write_model(PNModel(PNSpace(SD), 4, 4, inter), outdir, "sd44", list(range(0, 17, 2)))


def cpu_baseline(mf, seconds: float = 10.0):
    """REAL*4 reference loop nest (oracle/refcpu) on all host cores, bounded sample."""
    from oracle.refcpu import RefCPU
    threads = min(os.cpu_count() or 1, 16)
    inst = []
    for J2 in mf.J2s:
        r = RefCPU(*mf.args(J2), fp64=False, thresholds=True)
        r.set_threads(threads)
        inst.append((r, np.random.default_rng(J2).standard_normal(r.N)))
    for r, v in inst:  # warm-up: 3 H*v per instance
        for _ in range(3):
            r.apply(v)
    n = 0
    t0 = time.perf_counter()
    while True:
        for r, v in inst:
            r.apply(v)
            n += 1
        el = time.perf_counter() - t0
        if el >= seconds and n >= 20:
            break
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": n / el, "unit": "matvecs/s", "cores": threads, "kind": "port",
            "sample": f"{n} H*v over the nine sd 4p4n J=0..8 instances in {el:.1f} s "
                      f"(REAL*4 build with the reference's 1e-7 product thresholds)",
            "host_cpus": os.cpu_count(), "cpu_model": model}


# profile summary of the pn kernels of each small-block path
PN_PROFILE = {"stacked": "k_stk1_k_stk2", "staged": "k_pn_staged", "fused": "k_fuse_k_reduce",
              "stacked-mfma": "k_mx1_k_mx2", "fused-mfma": "k_mxf"}


def traffic_from_profile(path_name: str):
    """HBM-side bytes per launch of the pn kernels of the tuned path from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the
    copy-kernel factor recorded there), or None when that summary is absent."""
    name = PN_PROFILE.get(path_name, "")
    path = os.path.join(ROOT, "profiles", f"{name}_sd44_J0-8_pmc.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch"):
                vals["t"] = float(line.split("=")[1])
    return vals.get("t")


def time_to_5(log, train_samples: int = 256, epochs: int = 200, repeats: int = 3):
    """Wall time to the lowest 5 eigenvalues of every J = 0..8 instance of a
    held-out sd 4p4n interaction, to residual 1e-4 MeV (which bounds each
    eigenvalue's error by 1e-4 MeV), plain vs batched/mixed/warm-started.

    plain: one FP64 handle per J, solved one after another from the random start.
    fast: one batch of all J, FP32 or BF16 hi/lo H*v, warm start predicted by the
    network (inference and the fallback test inside the timed region).
    The FP64 oracle is the same solver at residual 1e-9."""
    from nushellx_amd.hv import Batch, Hamiltonian
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import PNModel, SideCache, write_model
    from nushellx_amd.warm import WarmStart, build_training_set

    J2s = list(range(0, 17, 2))
    space, base = PNSpace(SD), random_interaction(SD, seed=SEED)
    work = tempfile.mkdtemp(prefix="nsx_t5_")
    cache = SideCache()
    t0 = time.perf_counter()
    F, T, _, _ = build_training_set(space, 4, 4, J2s, base, range(1000, 1000 + train_samples), work,
                                    cache=cache, log=log)
    t_data = time.perf_counter() - t0
    ws = WarmStart(F.shape[1], T.shape[1], seed=0)
    t0 = time.perf_counter()
    losses = ws.fit(F, T, epochs=epochs, batch=16, lr=1e-3)
    t_fit = time.perf_counter() - t0
    log(f"warm start trained on {len(F)} samples ({t_data:.0f} s solves, {t_fit:.1f} s fit)")

    held = base.perturbed(HELD_OUT_SEED)
    mf = write_model(PNModel(space, 4, 4, held), work, "ws", J2s, cache=cache, densities=False, xtra="y")
    feats = held.parameter_vector()[None, :]
    hams64 = [Hamiltonian(*mf.args(J2)) for J2 in J2s]
    oracle = np.array([h.lanczos(nev=5, tol=1e-9)["evals"] for h in hams64])
    for h in hams64:
        h.autotune(iters=5)

    def best(fn):
        out = None
        for _ in range(repeats):
            t = time.perf_counter()
            r = fn()
            el = time.perf_counter() - t
            if out is None or el < out[0]:
                out = (el, r)
        return out

    t_plain, rp = best(lambda: [h.lanczos(nev=5, tol=1e-4) for h in hams64])
    ev_plain = np.array([r["evals"] for r in rp])
    res = {"workload": "sd 4p4n J=0..8, held-out interaction (base seed %d perturbed with seed %d)"
                       % (SEED, HELD_OUT_SEED),
           "criterion": "residual <= 1e-4 MeV for each of the 5 lowest pairs of every J",
           "plain_fp64_s": t_plain, "plain_matvecs": int(sum(r["matvecs"] for r in rp)),
           "plain_max_err_MeV": float(np.abs(ev_plain - oracle).max()),
           "train": {"samples": int(len(F)), "epochs": epochs, "solve_s": t_data, "fit_s": t_fit,
                     "loss_first": float(losses[0]), "loss_last": float(losses[-1])}}
    for prec in ("fp64", "fp32", "bf16x3"):
        hams = [Hamiltonian(*mf.args(J2), precision=prec) for J2 in J2s]
        b = Batch(hams, precision=prec)
        b.autotune(iters=5)
        t_rand, rr = best(lambda: b.lanczos(nev=5, tol=1e-4))

        def warm():
            v, fell = ws.start_vector(b, feats)
            r = b.lanczos(nev=5, tol=1e-4, start=v)
            r["fell"] = fell
            return r
        t_warm, rw = best(warm)
        res[prec] = {"batched_random_s": t_rand, "batched_random_matvecs": rr["matvecs"],
                     "batched_warm_s": t_warm, "batched_warm_matvecs": rw["matvecs"],
                     "fallback_instances": rw["fell"],
                     "max_err_random_MeV": float(np.abs(rr["evals"] - oracle).max()),
                     "max_err_warm_MeV": float(np.abs(rw["evals"] - oracle).max()),
                     "speedup_vs_plain": t_plain / t_warm}
        b.close()
        for h in hams:
            h.close()
    ok = {p: res[p]["max_err_warm_MeV"] <= 1e-4 for p in ("fp64", "fp32", "bf16x3")}
    fastest = min((p for p in ok if ok[p]), key=lambda p: res[p]["batched_warm_s"], default=None)
    res["fast"] = fastest
    res["speedup"] = res[fastest]["speedup_vs_plain"] if fastest else None
    return res


def ensemble(log, samples: int = 256):
    """Batched ensemble: `samples` perturbed interactions (V_0 + N(0, (0.1 |V_0|)^2))
    x J = 0, 2, 4 x sd 4p4n, 5p5n, 6p6n, all instances in one lockstep batch,
    lowest 5 states each to residual 1e-4 MeV; aggregate eigenpairs/s.  A linear
    ensemble: one template per (nucleus, J) from the base interaction's files,
    every sample's operators made on the device from its parameters."""
    from nushellx_amd.ensemble import ensemble_templates, linear_ensemble
    from nushellx_amd.hv import Batch, DeviceBuffer
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    J2s, nuclei = [0, 4, 8], [(4, 4), (5, 5), (6, 6)]
    space, base = PNSpace(SD), random_interaction(SD, seed=SEED)
    seeds = range(1000, 1000 + samples)
    templates = ensemble_templates(space, nuclei, J2s, base, tempfile.mkdtemp(prefix="nsx_ens_"))
    tmpl = templates[0]
    out = {"workload": f"{samples} samples x J=0,2,4 x sd 4p4n/5p5n/6p6n, lowest 5 each, residual 1e-4 MeV",
           "instances": samples * len(tmpl), "mode": "linear ensemble (operators made on the device)",
           "template_files_s": templates[2]["template_files_s"]}
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()  # tuning: the templates' own batch; its stacked shape is used
        small = Batch(tmpl, precision=prec)
        tuned = small.autotune(iters=5)
        small.close()
        t_tune = time.perf_counter() - t0
        # a stacked shape (VALU or matrix cores) carries over to the ensemble
        cfg = tuned if tuned.get("stacked") in (1, 3, 4) else None
        b, _, info = linear_ensemble(space, nuclei, J2s, base, seeds, "", precision=prec, config=cfg,
                                     templates=templates, log=log)
        setup = info["templates_s"] + info["theta_s"] + info["batch_s"]
        t0 = time.perf_counter()
        r = b.lanczos(nev=5, tol=1e-4)
        t_solve = time.perf_counter() - t0
        flops = b.work()[0]
        din, dout = DeviceBuffer(8 * b.N), DeviceBuffer(8 * b.N)
        din.upload(np.random.default_rng(0).standard_normal(b.N))
        ms_hv, ms_pn, _ = b.time(din, dout, warmup=3, iters=20)
        din.free()
        dout.free()
        out["N_total"] = int(b.N)
        n_inst = len(tmpl) * samples
        out[prec] = {"eigenpairs_per_s": n_inst * 5 / t_solve, "solve_s": t_solve, "matvecs": r["matvecs"],
                     "converged": int((r["converged_at"] >= 0).sum()), "setup_s": setup,
                     "templates_s": info["templates_s"], "theta_s": info["theta_s"], "batch_s": info["batch_s"],
                     "tune_s": t_tune, "host_s": r["host_seconds"], "hv_ms": ms_hv,
                     "hv_TFLOPs": flops / (ms_hv * 1e-3) / 1e12, "tuned": cfg}
        log(f"ensemble {prec}: {n_inst * 5 / t_solve:.0f} eigenpairs/s, setup {setup:.2f} s, solve {t_solve:.2f} s")
        b.close()
    for h in tmpl:
        h.close()
    return out


def precision_study(log):
    """sd shell 6p6n, J = 0, 2, 4 (N = 3372, 13562, 15089; same seeded
    interaction as the headline workload): the lowest five states of every J
    with the block contractions in FP64, FP32 and BF16 hi/lo (the recurrence,
    reorthogonalisation and Ritz solve stay FP64), to residual 1e-6 MeV, in one
    batch per precision.  Eigenvalue errors are against the FP64 solve at
    residual 1e-9 MeV (the solver that the tests pin to dense eigvalsh)."""
    from nushellx_amd.hv import Batch, Hamiltonian
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import SD, PNSpace
    from nushellx_amd.synth.writer import PNModel, write_model
    work = tempfile.mkdtemp(prefix="nsx_prec_")
    J2s = [0, 4, 8]
    mf = write_model(PNModel(PNSpace(SD), 6, 6, random_interaction(SD, seed=SEED)), work, "sd66", J2s)
    hams = [Hamiltonian(*mf.args(J2)) for J2 in J2s]
    ref_b = Batch(hams)
    ref = ref_b.lanczos(nev=5, tol=1e-9)["evals"]
    ref_b.close()
    out = {"workload": "sd 6p6n J=0,2,4 (N = %s), lowest 5 each, residual 1e-6 MeV" % "/".join(str(h.N) for h in hams),
           "reference": "FP64 solve at residual 1e-9 MeV", "evals_MeV": ref.tolist()}
    for prec in ("fp64", "fp32", "bf16x3"):
        b = Batch(hams, precision=prec)
        b.autotune(iters=5)
        b.lanczos(nev=5, tol=1e-6)  # first solve captures the graphs
        t0 = time.perf_counter()
        r = b.lanczos(nev=5, tol=1e-6)
        el = time.perf_counter() - t0
        err = np.abs(r["evals"] - ref)
        out[prec] = {"s": el, "matvecs": int(r["matvecs"]), "max_err_MeV": float(err.max()),
                     "err_MeV": err.tolist()}
        log(f"precision study {prec}: {el * 1e3:.1f} ms, {r['matvecs']} H*v, max eigenvalue error {err.max():.2e} MeV")
        b.close()
    for h in hams:
        h.close()
    return out


def pf_warm(log, train_samples: int, nprot: int = 4, nneut: int = 4, chunk: int = 256, held_n: int = 8,
            threads: int = 16, train_tol: float = 1e-5):
    """Learned warm start on pf 4p4n J=0 (N = 41,355; 195 TBMEs ~ N(0, 1.5 MeV),
    SPEs ~ U[-10, -4] MeV), or pf 6p6n J=0 (N = 1,777,116) with nprot = nneut
    = 6: network trained on the lowest-8 Lanczos eigenvectors of
    `train_samples` perturbed interactions (solved `chunk` at a time), then
    `held_n` held-out interactions solved in one lockstep batch from the random
    and from the predicted start (lowest 5, residual 1e-4 MeV)."""
    from nushellx_amd.ensemble import build_instances, solve_ensemble, write_ensemble
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import PF, PNSpace
    from nushellx_amd.warm import WarmStart, build_training_set_batched
    base = random_interaction(PF, seed=SEED, spe_range=(-10.0, -4.0))
    space, J2s = PNSpace(PF), [0]
    work = tempfile.mkdtemp(prefix="nsx_pfw_")
    t0 = time.perf_counter()
    F, T, _, _ = build_training_set_batched(space, nprot, nneut, J2s, base, range(1000, 1000 + train_samples), work,
                                            tol=train_tol, chunk=chunk, threads=threads, log=log)
    t_data = time.perf_counter() - t0
    ws = WarmStart(F.shape[1], T.shape[1], seed=0)
    t0 = time.perf_counter()
    losses = ws.fit(F, T, epochs=200, batch=16, lr=1e-3)
    t_fit = time.perf_counter() - t0
    log(f"warm start fitted ({t_fit:.1f} s)")
    held = list(range(9001, 9001 + held_n))
    hams = build_instances(write_ensemble(space, [(nprot, nneut)], J2s, base, held, work + "/held"), J2s,
                           threads=threads)
    feats = np.array([base.perturbed(s).parameter_vector() for s in held])
    r0 = solve_ensemble(hams, nev=5, tol=1e-4)
    b = r0["batch"]
    t0 = time.perf_counter()
    v, fell = ws.start_vector(b, feats)
    rw = b.lanczos(nev=5, tol=1e-4, start=v)
    t_warm = time.perf_counter() - t0
    t0 = time.perf_counter()
    rr = b.lanczos(nev=5, tol=1e-4)
    t_rand = time.perf_counter() - t0
    out = {"workload": f"pf {nprot}p{nneut}n J=0, {len(F)} training samples, {held_n} held-out interactions "
                       f"in one batch",
           "train_solve_s": t_data, "fit_s": t_fit, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
           "random": {"s": t_rand, "matvecs": rr["matvecs"], "converged_at": rr["converged_at"].tolist()},
           "warm": {"s": t_warm, "matvecs": rw["matvecs"], "converged_at": rw["converged_at"].tolist(),
                    "fallback": fell},
           "max_eval_diff_MeV": float(np.abs(rw["evals"] - rr["evals"]).max())}
    b.close()
    return out


def host_f32(log, hams, iters: int = 50):
    """The drop-in as NuTRL calls it: nsx_hv_apply_f32 on host REAL*4 vectors
    (coefR in, coefL out; PCIe transfers, device widening and narrowing
    included), per instance, one instance after another."""
    per = []
    for h in hams:
        v = np.random.default_rng(h.N).standard_normal(h.N).astype(np.float32)
        h.multiply(v)
        t0 = time.perf_counter()
        for _ in range(iters):
            h.multiply(v)
        per.append(1e6 * (time.perf_counter() - t0) / iters)
    out = {"workload": "sd 4p4n J=0..8, one REAL*4 host-vector H*v per instance in turn",
           "us_per_instance": per, "matvecs_per_s": len(hams) / (1e-6 * sum(per))}
    log(f"host REAL*4 path: {out['matvecs_per_s']:.0f} H*v/s ({min(per):.0f}-{max(per):.0f} us per instance)")
    return out


def deterministic(log, hams, ms_default):
    """The headline batch planned for a bitwise reproducible H*v (the fused
    column path: X and Y in LDS, column partials summed in plan order): its
    H*v time beside the default path's, and three H*v compared bit for bit."""
    from nushellx_amd.hv import Batch, DeviceBuffer, device_sync
    os.environ["NSX_DETERMINISTIC"] = "1"
    try:
        b = Batch(hams)
    finally:
        os.environ.pop("NSX_DETERMINISTIC", None)
    v = np.random.default_rng(3).standard_normal(b.N)
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    ms_hv, ms_pn, ms_red = b.time(din, dout, warmup=10, iters=200)
    outs = []
    for _ in range(3):
        b.multiply_dev(din, dout)
        device_sync()
        w = np.empty_like(v)
        dout.download(w)
        outs.append(w)
    out = {"path": b.path(), "hv_ms": ms_hv, "pn_ms": ms_pn, "reduce_ms": ms_red,
           "cost_vs_default": ms_hv / ms_default,
           "bitwise_identical": bool(all(np.array_equal(o, outs[0]) for o in outs[1:]))}
    log(f"deterministic: {out['path']} H*v {ms_hv * 1e3:.1f} us ({out['cost_vs_default']:.2f}x the default), "
        f"bitwise identical {out['bitwise_identical']}")
    din.free()
    dout.free()
    b.close()
    return out


def separable(log, mf, cuts=((0.0, 0.0), (0.3, 0.05), (1.0, 0.1))):
    """NuTRLZ's separable pn operator on the same workload: the .dop files NuDiop
    would write at several cuts (eigenvalue / sqrt(2lam+1) in MeV, amplitude), the
    batched H*v time of each (autotuned like the headline batch), and how far the
    lowest five J=0 eigenvalues move from the full operator's."""
    from nushellx_amd.hv import Batch, DeviceBuffer, Hamiltonian
    from nushellx_amd.synth.diop import diop_sections, read_oph, write_dop
    items, max_lmo = read_oph(mf.intfile + ".oph")
    ref = Hamiltonian(*mf.args(0)).lanczos(nev=5, tol=1e-8)["evals"]
    out = []
    for k, (cut, mcut) in enumerate(cuts):
        secs = diop_sections(items, max_lmo, cut, mcut)
        dop = mf.intfile + f"d{k}"
        write_dop(dop + ".dop", secs, max_lmo)
        hams = [Hamiltonian(*mf.args(J2), dopfile=dop) for J2 in mf.J2s]
        b = Batch(hams)
        tuned = b.autotune(iters=5)
        flops, flops_like, _ = b.work()
        v = np.random.default_rng(2).standard_normal(b.N)
        din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
        ms_hv, ms_pn, _ = b.time(din, dout, warmup=10, iters=500)
        ev = hams[0].lanczos(nev=5, tol=1e-8)["evals"]
        row = {"cut_MeV": cut, "mcut": mcut, "terms": sum(len(t) for _, _, t in secs),
               "hv_ms": ms_hv, "pn_ms": ms_pn, "matvecs_per_s": len(hams) / (ms_hv * 1e-3),
               "flops_pn": flops - flops_like, "pn_TFLOPs": (flops - flops_like) / (ms_pn * 1e-3) / 1e12,
               "J0_eval_shift_MeV": float(np.abs(ev - ref).max()), "stacked": int(tuned.get("stacked", 0))}
        log(f"separable cut {cut}/{mcut}: {row['terms']} terms, H*v {ms_hv * 1e3:.1f} us, "
            f"J=0 eigenvalue shift {row['J0_eval_shift_MeV']:.3g} MeV")
        out.append(row)
        din.free()
        dout.free()
        b.close()
    return out


def large(log, steps: int = 5, cpu: bool = True):
    """One large instance: pf shell (0f7/2 1p3/2 0f5/2 1p1/2), 6 protons + 6
    neutrons, J = 0 (N = 1,777,116; blocks up to 615 x 615), seeded interaction
    (TBMEs ~ N(0, 1.5 MeV), SPEs ~ U[-10, -4] MeV).  Its blocks run on the FP64
    matrix-core tiles.  Reports the FP64 H*v time, the reference's sparse flops
    per second, the time to the lowest five states (FP64, and with the BF16
    hi/lo contractions), and one H*v of the CPU restatement on the host cores."""
    from nushellx_amd.hv import DeviceBuffer, Hamiltonian, device_sync
    from nushellx_amd.synth.interaction import random_interaction
    from nushellx_amd.synth.mspace import PF, PNSpace
    from nushellx_amd.synth.writer import PNModel, write_model
    work = tempfile.mkdtemp(prefix="nsx_pf66_")
    t0 = time.perf_counter()
    log("large: writing pf 6p6n J=0 operator files")
    mf = write_model(PNModel(PNSpace(PF), 6, 6, random_interaction(PF, seed=SEED, spe_range=(-10.0, -4.0))),
                     work, "pf66", [0])
    t_files = time.perf_counter() - t0
    t0 = time.perf_counter()
    h = Hamiltonian(*mf.args(0))
    t_setup = time.perf_counter() - t0
    flops, flops_like, _ = h.work()
    v = np.random.default_rng(1).standard_normal(h.N)
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    h.multiply_dev(din, dout)
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.multiply_dev(din, dout)
    device_sync()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    out = {"workload": "pf 6p6n J=0, N=%d, one FP64 H*v" % h.N, "files_s": t_files, "setup_s": t_setup,
           "hv_ms": ms, "flops_ref": flops, "TFLOPs_ref": flops / (ms * 1e-3) / 1e12}
    v32 = v.astype(np.float32)
    h.multiply(v32)
    t0 = time.perf_counter()
    for _ in range(steps):
        h.multiply(v32)
    out["host_f32_ms"] = 1e3 * (time.perf_counter() - t0) / steps  # REAL*4 host vectors, PCIe included
    log(f"large: pf 6p6n H*v {ms:.1f} ms ({out['TFLOPs_ref']:.2f} TFLOP/s of the reference's sparse work)")
    din.free()
    dout.free()
    # the lowest five J=0 states to residual 1e-4 MeV, FP64 thick restart from the random start
    t0 = time.perf_counter()
    r = h.lanczos(nev=5, tol=1e-4)
    out["lanczos"] = {"seconds": time.perf_counter() - t0, "matvecs": int(r["matvecs"]),
                      "restarts": int(r["restarts"]), "converged": int(r["converged"]),
                      "evals": [float(e) for e in r["evals"]], "max_residual": float(np.max(r["residuals"]))}
    log(f"large: lowest 5 to 1e-4 MeV in {out['lanczos']['seconds']:.2f} s, {r['matvecs']} H*v")
    # the same solve with the BF16 hi/lo contractions (BF16 matrix cores), Lanczos in FP64
    hb = Hamiltonian(*mf.args(0), precision="bf16x3")
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    hb.multiply_dev(din, dout)
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        hb.multiply_dev(din, dout)
    device_sync()
    ms_b = 1e3 * (time.perf_counter() - t0) / steps
    din.free()
    dout.free()
    t0 = time.perf_counter()
    rb = hb.lanczos(nev=5, tol=1e-4)
    out["bf16x3"] = {"hv_ms": ms_b, "lanczos_s": time.perf_counter() - t0, "matvecs": int(rb["matvecs"]),
                     "max_eval_diff_vs_fp64_MeV": float(np.abs(rb["evals"] - r["evals"]).max())}
    out["bf16x3"]["speedup_vs_fp64"] = out["lanczos"]["seconds"] / out["bf16x3"]["lanczos_s"]
    log(f"large: BF16 hi/lo H*v {ms_b:.2f} ms, lowest 5 in {out['bf16x3']['lanczos_s']:.2f} s "
        f"({out['bf16x3']['speedup_vs_fp64']:.2f}x FP64), eigenvalues within "
        f"{out['bf16x3']['max_eval_diff_vs_fp64_MeV']:.1e} MeV of the FP64 solve")
    hb.close()
    # deterministic FP64 plan: one stage-2 workgroup per large W tile
    os.environ["NSX_DETERMINISTIC"] = "1"
    try:
        hd = Hamiltonian(*mf.args(0))
    finally:
        os.environ.pop("NSX_DETERMINISTIC", None)
    din, dout = DeviceBuffer.from_array(v), DeviceBuffer(v.nbytes)
    hd.multiply_dev(din, dout)
    device_sync()
    w1 = np.empty_like(v)
    dout.download(w1)
    t0 = time.perf_counter()
    for _ in range(steps):
        hd.multiply_dev(din, dout)
    device_sync()
    ms_d = 1e3 * (time.perf_counter() - t0) / steps
    w2 = np.empty_like(v)
    dout.download(w2)
    out["deterministic"] = {"hv_ms": ms_d, "cost_vs_default": ms_d / ms, "bitwise_identical": bool(np.array_equal(w1, w2))}
    log(f"large: deterministic FP64 H*v {ms_d:.2f} ms ({ms_d / ms:.2f}x), bitwise identical "
        f"{out['deterministic']['bitwise_identical']}")
    din.free()
    dout.free()
    hd.close()
    if cpu:
        from oracle.refcpu import RefCPU
        r = RefCPU(*mf.args(0), fp64=False, thresholds=True)
        threads = min(os.cpu_count() or 1, 16)
        r.set_threads(threads)
        t0 = time.perf_counter()
        r.apply(v)
        cpu_s = time.perf_counter() - t0
        out["cpu_port_s"] = cpu_s
        out["cpu_threads"] = threads
        out["gpu_over_cpu"] = cpu_s / (ms * 1e-3)
        out["lanczos"]["cpu_est_s"] = cpu_s * out["lanczos"]["matvecs"]  # H*v only
        log(f"large: CPU restatement {cpu_s:.2f} s per H*v on {threads} threads")
    h.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-t5", action="store_true", help="skip the time-to-5-eigenvalues comparison")
    ap.add_argument("--train-samples", type=int, default=256)
    ap.add_argument("--copy-mb", type=int, default=0, help="also run the streaming copy kernel (MiB)")
    ap.add_argument("--ensemble-samples", type=int, default=256,
                    help="interaction samples of the batched ensemble (0 skips it)")
    ap.add_argument("--no-sep", action="store_true", help="skip the separable-operator (NuTRLZ) timings")
    ap.add_argument("--no-prec", action="store_true", help="skip the sd 6p6n precision study")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the pf 6p6n J=0 H*v (N = 1.78 M; about a minute, mostly writing its files)")
    ap.add_argument("--pf66-warm-samples", type=int, default=0,
                    help="training samples of the pf 6p6n J=0 warm-start study (0 skips it; 16 take ~6 min)")
    ap.add_argument("--pf-warm-samples", type=int, default=2000,
                    help="training samples of the pf 4p4n warm-start study (0 skips it; solved 256 at a time "
                         "as linear ensembles)")
    ap.add_argument("--no-det", action="store_true", help="skip the deterministic-mode timings")
    ap.add_argument("--shape", default="",
                    help='fixed plan shape for the headline batch instead of autotuning, as JSON, e.g. '
                         '\'{"stacked": 1, "stk_cap": 2048, "stk_work": 256}\'')
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent J/sample instances would be sharded round-robin across GPUs).",
              file=sys.stderr)
        sys.exit(2)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    from nushellx_amd.hv import Batch, DeviceBuffer, Hamiltonian, device_sync
    from nushellx_amd._lib import device_info, stream_copy

    tmp = tempfile.mkdtemp(prefix="nsx_bench_")
    mf = make_workload(tmp)
    hams = [Hamiltonian(*mf.args(J2)) for J2 in mf.J2s]
    if args.shape:  # a fixed plan shape instead of the autotuner (reproducible profiles)
        tuned = dict({"fat_cap": 5120, "max_phases": 6, "threads": 256, "mode": 0, "atomic": 1, "stacked": 1},
                     **json.loads(args.shape))
        batch = Batch(hams, config=tuned)
        log(f"fixed shape: {tuned}")
    else:
        batch = Batch(hams)
        tuned = batch.autotune(iters=10)   # plan shape and kernel choice, chosen before timing
        log(f"autotuned: {tuned}")
    flops, flops_like, bytes_alg = batch.work()
    v = np.random.default_rng(1).standard_normal(batch.N)
    din = DeviceBuffer.from_array(v)
    dout = DeviceBuffer(v.nbytes)

    for _ in range(args.warmup):
        batch.multiply_dev(din, dout)
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.multiply_dev(din, dout)
    device_sync()
    el = time.perf_counter() - t0
    ms_step = 1e3 * el / args.steps
    value = len(hams) * args.steps / el
    log(f"timed {args.steps} batched H*v: {value:.0f} matvecs/s")

    # HIP events on the batch stream around loops of back-to-back launches: the
    # whole H*v, and its pn kernels alone (no per-launch event overhead)
    ms_hv, ms_pn, ms_red = batch.time(din, dout, warmup=10, iters=min(args.steps, 1000))
    path = batch.path()
    pname = path.split()[0]
    stacked = pname.startswith("stacked")
    # the stacked stage-1 launch (VALU or matrix cores) and the fused
    # matrix-core path with "+like" also carry the like-particle blocks, so their
    # pn launches are the whole H*v; the other paths do that part in the reduce
    flops_pn = flops if stacked or "+like" in path else flops - flops_like
    achieved = flops_pn / (ms_pn * 1e-3) / 1e12
    traffic = traffic_from_profile(pname)
    kernel = {"stacked": "k_stk1 + k_stk2 (stacked two-stage pn path with the like-particle blocks, FP64 VALU)",
              "staged": "k_pn_staged (pn two-body part, FP64 VALU)",
              "fused": "k_fuse (fused column path, pn two-body part, FP64 VALU)",
              "stacked-mfma": "k_mx1 + k_mx2 (stacked stages on FP64 16x16x4 matrix-core tiles, with the "
                              "like-particle blocks)",
              "fused-mfma": "k_mxf (pn two-body part: both stacked stages in one kernel on FP64 16x16x4 "
                            "matrix-core tiles, X and Y in registers)"}[pname]
    mfma = pname.endswith("-mfma")
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                "engine": ("FP64 matrix cores (v_mfma_f64_16x16x4f64); the tiles are bound by operand-load "
                           "latency, not by the matrix-core peak" if mfma else
                           "FP64 vector FMA (VALU), whose peak equals the FP64 matrix-core peak on gfx950; "
                           "the kernels are bound by load latency and instruction issue, not by either peak"),
                "kernel": kernel, "kernel_ms": ms_pn, "flops_per_launch": flops_pn,
                "hbm_GBs_alg": (bytes_alg / (ms_hv * 1e-3)) / 1e9,
                "hbm_GBs_counters": (traffic / (ms_pn * 1e-3)) / 1e9 if traffic else None}

    out = {
        "metric": "time-to-5-eigenvalues @1e-4 MeV & matvecs/s: mixed+warm vs plain FP64",
        "value": value,
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sd-shell interaction; NuShellX-format operator files generated before timing)",
        "config": {"workload": "sd 4p4n J=0..8 plain FP64 H*v, 9 instances batched per step",
                   "N_total": int(batch.N), "instances": len(hams),
                   "flops_per_step": flops, "bytes_per_step": bytes_alg},
        "roofline": roofline,
        "kernel_ms": {"hv": ms_hv, "pn": ms_pn, "reduce": ms_red},
        "tuned": tuned,
        "path": path,
    }
    try:
        out["device"] = list(device_info())
    except Exception:
        pass
    if args.copy_mb:
        nb = args.copy_mb << 20
        ms = stream_copy(nb, iters=5)
        out["copy_kernel"] = {"bytes": nb, "ms": ms, "GBs": 2 * nb / (ms * 1e-3) / 1e9}
    din.free()
    dout.free()
    batch.close()
    out["host_f32"] = host_f32(log, hams)
    if not args.no_det:
        out["deterministic"] = deterministic(log, hams, ms_hv)
    if not args.no_t5:
        out["time_to_5"] = time_to_5(log, train_samples=args.train_samples)
    if not args.no_sep:
        out["separable"] = separable(log, mf)
    if not args.no_prec:
        out["precision"] = precision_study(log)
    if not args.no_large:
        try:
            out["large"] = large(log, cpu=not args.no_cpu)
        except Exception as e:  # the headline line must print whatever happens here
            out["large"] = {"error": repr(e)}
            log(f"large: failed: {e!r}")
    if args.ensemble_samples > 0:
        out["ensemble"] = ensemble(log, args.ensemble_samples)
    if args.pf_warm_samples > 0:
        out["pf_warm"] = pf_warm(log, args.pf_warm_samples)
    if args.pf66_warm_samples > 0:
        out["pf66_warm"] = pf_warm(log, args.pf66_warm_samples, nprot=6, nneut=6, chunk=4, held_n=4, threads=4,
                                   train_tol=1e-5)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mf, args.cpu_seconds)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        if "time_to_5" in out:
            per = out["cpu_baseline"]["value"]
            out["cpu_baseline"]["time_to_5_est_s"] = out["time_to_5"]["plain_matvecs"] / per
    print(json.dumps(out))


if __name__ == "__main__":
    main()
