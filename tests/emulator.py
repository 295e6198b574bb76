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
                                                         for h in ("nsx_model.hpp", "nsx_plan.hpp", "nsx_io.hpp", "nsx_eig.hpp")])
        if not os.path.exists(SO) or os.path.getmtime(SO) < newest:
            subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", SO, *SRCS], check=True)
        L = C.CDLL(SO)
        L.emu_hv.argtypes = [C.c_char_p] * 4 + [C.c_void_p, C.c_void_p, C.c_int64]
        L.emu_hv_sep.argtypes = [C.c_char_p] * 5 + [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_double)]
        L.emu_hv_big.argtypes = [C.c_char_p] * 5 + [C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.emu_plan_stats.argtypes = [C.c_char_p] * 4 + [C.c_void_p]
        L.emu_last_error.restype = C.c_char_p
        L.emu_sixj.argtypes = [C.c_int] * 6
        L.emu_sixj.restype = C.c_double
        L.emu_hv_batch.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_int, C.c_void_p, C.c_void_p, C.c_int64,
                                   C.c_void_p]
        L.emu_hv_stacked.argtypes = [C.c_char_p] * 4 + [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int]
        L.emu_sym_eig.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.emu_hv_fused.argtypes = [C.c_char_p] * 4 + [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_int,
                                                      C.c_void_p]
        L.emu_hv_mstk.argtypes = [C.c_char_p] * 4 + [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_int,
                                                     C.c_void_p]
        L.emu_hv_mxf.argtypes = [C.c_char_p] * 4 + [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_int,
                                                    C.c_int, C.c_void_p]
        _LIB = L
    return _LIB


def emu_hv(args, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    rc = lib().emu_hv(*[a.encode() for a in args], v.ctypes.data, out.ctypes.data, len(v))
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out


def emu_hv_sep(args, dop, v):
    """H*v of the separable model (<dop>.dop) through the device plan; returns (w, flops)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    f = C.c_double()
    rc = lib().emu_hv_sep(*[a.encode() for a in args], dop.encode(), v.ctypes.data, out.ctypes.data, len(v),
                          C.byref(f))
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out, f.value


def emu_hv_big(args, v, big_min, dop=None):
    """H*v with the items above big_min on the large-block GEMM plan; returns
    (w, (stage-1 tiles, stage-2 tiles, X references, stage-2 terms))."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    tiles = np.zeros(4, dtype=np.int64)
    rc = lib().emu_hv_big(*[a.encode() for a in args], dop.encode() if dop else None, big_min, v.ctypes.data,
                          out.ctypes.data, len(v), tiles.ctypes.data)
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out, tuple(int(x) for x in tiles)


def plan_stats(args):
    out = np.zeros(8, dtype=np.int64)
    rc = lib().emu_plan_stats(*[a.encode() for a in args], out.ctypes.data)
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return dict(zip(["items2", "phases", "pairs", "items", "blocks", "part_size", "lds_buf", "operator_doubles"], out.tolist()))


def sixj(*a):
    return lib().emu_sixj(*a)


def sym_eig(A, full=True):
    """Host eigensolver of the Lanczos driver (nsx_eig.hpp): (w, Z) or (w, last row)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    Z = np.zeros((n, n) if full else n)
    if lib().emu_sym_eig(A.ctypes.data, n, int(full), w.ctypes.data, Z.ctypes.data):
        raise RuntimeError("QL did not converge")
    return w, (Z.T if full else Z)  # Z stored column-major: rows of the C array are eigenvectors


def emu_hv_batch(arg_lists, v, merge):
    """Emulated H*v of a batch planned in one pass (merge=False) or per
    instance and merged (merge=True); returns (w, store sizes Ad/Bh/H)."""
    flat = [a.encode() for args in arg_lists for a in args]
    arr = (C.c_char_p * len(flat))(*flat)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    sizes = np.zeros(3, dtype=np.int64)
    rc = lib().emu_hv_batch(len(arg_lists), arr, int(merge), v.ctypes.data, out.ctypes.data, len(v),
                            sizes.ctypes.data)
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out, sizes


def emu_hv_fused(args, v, cap=3072, work=250000, threads=256):
    """Emulated H*v through the fused column plan; returns (W, stats) with stats =
    (tasks, rounds, largest round LDS, partial doubles, largest task units)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    st = np.zeros(5, dtype=np.int64)
    rc = lib().emu_hv_fused(*[a.encode() for a in args], v.ctypes.data, out.ctypes.data, len(v), cap, work, threads,
                            st.ctypes.data)
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out, st


def emu_hv_stacked(args, v, cap=2048, work=512):
    """Emulated H*v through the stacked two-stage plan."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    rc = lib().emu_hv_stacked(*[a.encode() for a in args], v.ctypes.data, out.ctypes.data, len(v), cap, work)
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out


def emu_hv_mstk(args, v, units1=4, terms2=16, big_min=0):
    """Emulated H*v through the matrix-core stacked plan (lane-level model of
    k_mx1 / k_mx2); returns (w, [stage-1 tasks, stage-2 tasks, X size, W writes])."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    st = np.zeros(4, dtype=np.int64)
    rc = lib().emu_hv_mstk(*[a.encode() for a in args], v.ctypes.data, out.ctypes.data, len(v), units1, terms2,
                           big_min, st.ctypes.data)
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out, st


def emu_hv_mxf(args, v, steps=2, slots=4, like=1, det=0):
    """Emulated H*v through the fused matrix-core plan (lane-level model of
    k_mxf); returns (w, [tasks, steps, products, largest inner size in K steps
    + 100 when the like-particle part is folded into the tasks])."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty_like(v)
    st = np.zeros(4, dtype=np.int64)
    rc = lib().emu_hv_mxf(*[a.encode() for a in args], v.ctypes.data, out.ctypes.data, len(v), steps, slots, like,
                          det, st.ctypes.data)
    if rc:
        raise RuntimeError(lib().emu_last_error().decode())
    return out, st
