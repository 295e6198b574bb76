"""Build and load the CPU plan emulator (test helper, tests/plan_emulator.cpp)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SO = os.path.join(HERE, "_plan_emulator.so")
SRCS = [os.path.join(HERE, "plan_emulator.cpp")] + [
    os.path.join(ROOT, "nushellx_amd", "csrc", f) for f in ("nsx_io.cpp", "nsx_model.cpp", "nsx_plan.cpp")]

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        newest = max(os.path.getmtime(s) for s in SRCS + [os.path.join(ROOT, "nushellx_amd", "csrc", h)
                                                         for h in ("nsx_model.hpp", "nsx_plan.hpp", "nsx_io.hpp")])
        if not os.path.exists(SO) or os.path.getmtime(SO) < newest:
            subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", SO, *SRCS], check=True)
        L = C.CDLL(SO)
        L.emu_hv.argtypes = [C.c_char_p] * 4 + [C.c_void_p, C.c_void_p, C.c_int64]
        L.emu_plan_stats.argtypes = [C.c_char_p] * 4 + [C.c_void_p]
        L.emu_last_error.restype = C.c_char_p
        L.emu_sixj.argtypes = [C.c_int] * 6
        L.emu_sixj.restype = C.c_double
        _LIB = L
    return _LIB


def emu_hv(args, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    rc = lib().emu_hv(*[a.encode() for a in args], v.ctypes.data, out.ctypes.data, len(v))
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out


def plan_stats(args):
    out = np.zeros(8, dtype=np.int64)
    rc = lib().emu_plan_stats(*[a.encode() for a in args], out.ctypes.data)
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return dict(zip(["items2", "phases", "pairs", "items", "blocks", "part_size", "lds_buf", "operator_doubles"], out.tolist()))


def sixj(*a):
    return lib().emu_sixj(*a)
