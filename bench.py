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
    return {"value": n / el, "unit": "matvecs/s", "cores": threads, "kind": "port",
            "sample": f"{n} H*v over the nine sd 4p4n J=0..8 instances in {el:.1f} s "
                      f"(REAL*4 build with the reference's 1e-7 product thresholds)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent J/sample instances would be sharded round-robin across GPUs).",
              file=sys.stderr)
        sys.exit(2)

    from nushellx_amd.hv import Batch, DeviceBuffer, Hamiltonian, device_sync
    from nushellx_amd._lib import device_info

    tmp = tempfile.mkdtemp(prefix="nsx_bench_")
    mf = make_workload(tmp)
    hams = [Hamiltonian(*mf.args(J2)) for J2 in mf.J2s]
    batch = Batch(hams)
    tuned = batch.autotune(iters=10)   # staged-kernel shape, chosen before timing
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

    # HIP events on the batch stream around every launch of the timed loop
    ms_hv, ms_pn, ms_red = batch.time(din, dout, warmup=10, iters=min(args.steps, 1000))
    flops_pn = flops - flops_like
    achieved = flops_pn / (ms_pn * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                "kernel": "k_pn_partial", "kernel_ms": ms_pn, "flops_per_launch": flops_pn}

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
    }
    try:
        out["device"] = list(device_info())
    except Exception:
        pass
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mf, args.cpu_seconds)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
