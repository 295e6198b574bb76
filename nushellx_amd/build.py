"""Build the HIP library ``libnsx.so`` in-tree for gfx950 (hipcc, no JIT cache).

Every source is compiled to its own object in parallel (each kernel is
launched from the file that defines it, so no relocatable device code is
needed), then the objects are linked into the shared library."""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnsx.so")
SOURCES = ["nsx_io.cpp", "nsx_model.cpp", "nsx_plan.cpp", "nsx_kernels.hip", "nsx_lowp.hip",
           "nsx_trl.hip", "nsx_mlp.hip", "nsx_api.hip", "nsx_files.cpp", "nsx_xpct.hip", "nsx_nuvec.hip",
           "nsx_ens.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose: bool = False, csrc: str = CSRC, out: str = LIB, extra=()) -> str:
    srcs = [os.path.join(csrc, s) for s in SOURCES if os.path.exists(os.path.join(csrc, s))]
    with tempfile.TemporaryDirectory(prefix="nsx_build_") as tmp:
        objs = [os.path.join(tmp, os.path.basename(s) + ".o") for s in srcs]

        def compile_one(pair):
            src, obj = pair
            # .cpp sources hold host code only; they still go through hipcc for one toolchain
            lang = ["-x", "hip"] if src.endswith(".hip") else []
            return subprocess.run(["hipcc", *FLAGS, *extra, *lang, "-c", src, "-o", obj], capture_output=True,
                                  text=True)

        with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 4)) as ex:
            results = list(ex.map(compile_one, zip(srcs, objs)))
        failed = [r for r in results if r.returncode != 0]
        for r in results:
            if verbose or r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
        if failed:
            raise RuntimeError("hipcc failed building libnsx.so")
        tmp_lib = out + ".tmp"
        r = subprocess.run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib, *objs],
                           capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed linking libnsx.so")
    os.replace(tmp_lib, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
