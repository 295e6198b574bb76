"""ORACLE (test infrastructure only): ctypes binding of ``librefcpu.so``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "librefcpu.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librefcpu.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.refcpu_create.argtypes = [C.c_char_p] * 4 + [C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.refcpu_create_sep.argtypes = [C.c_char_p] * 5 + [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.refcpu_destroy.argtypes = [C.c_void_p]
        L.refcpu_dim.argtypes = [C.c_void_p]
        L.refcpu_dim.restype = C.c_int64
        L.refcpu_basd.argtypes = [C.c_void_p]
        L.refcpu_block.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.refcpu_lcoef.argtypes = [C.c_void_p, C.c_void_p]
        L.refcpu_trunc.argtypes = [C.c_void_p, C.c_void_p]
        L.refcpu_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.refcpu_work.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_void_p]
        L.refcpu_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.refcpu_last_error.restype = C.c_char_p
        L.refcpu_tfind.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.refcpu_sixj.argtypes = [C.c_int] * 6
        L.refcpu_sixj.restype = C.c_double
        _LIB = L
    return _LIB


class RefCPU:
    """The reference H*v restated in C; ``fp64=False`` is the REAL*4 build."""

    def __init__(self, nus_prefix, opfile_a, opfile_b, intfile, fp64=True, thresholds=False,
                 dopfile=None, limits=True):
        """``dopfile``: NuTRLZ's separable pn part from ``<dopfile>.dop``; ``limits``
        applies its zero limits (False keeps NuTRL's, to compare the algebra)."""
        L = lib()
        h = C.c_void_p()
        if dopfile is None:
            rc = L.refcpu_create(nus_prefix.encode(), opfile_a.encode(), opfile_b.encode(),
                                 intfile.encode(), int(fp64), int(thresholds), C.byref(h))
        else:
            rc = L.refcpu_create_sep(nus_prefix.encode(), opfile_a.encode(), opfile_b.encode(),
                                     intfile.encode(), dopfile.encode(), int(fp64), int(thresholds),
                                     int(limits), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"refcpu_create failed ({rc}): {L.refcpu_last_error().decode()}")
        self._h = h
        self.N = int(L.refcpu_dim(h))
        self.basd = int(L.refcpu_basd(h))

    def close(self):
        if self._h:
            lib().refcpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n: int):
        lib().refcpu_set_threads(self._h, n)

    def apply(self, v: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty_like(v)
        lib().refcpu_apply(self._h, v.ctypes.data, out.ctypes.data)
        return out

    def blocks(self):
        L = lib()
        out = []
        a, b, p, d = C.c_int(), C.c_int(), C.c_int64(), C.c_int64()
        for i in range(self.basd):
            L.refcpu_block(self._h, i, C.byref(a), C.byref(b), C.byref(p), C.byref(d))
            out.append((a.value, b.value, p.value, d.value))
        return out

    def lcoef(self) -> np.ndarray:
        m = np.zeros(self.N, dtype=np.uint8)
        lib().refcpu_lcoef(self._h, m.ctypes.data)
        return m.astype(bool)

    def trunc(self):
        """(truncated, mask) of the N-hbar-omega window (mask = lcoef when there is none)."""
        m = np.zeros(self.N, dtype=np.uint8)
        t = lib().refcpu_trunc(self._h, m.ctypes.data)
        return bool(t), m.astype(bool)

    def work(self):
        f = C.c_double()
        cnt = np.zeros(5, dtype=np.int64)
        lib().refcpu_work(self._h, C.byref(f), cnt.ctypes.data)
        return f.value, cnt

    def dense(self) -> np.ndarray:
        """The dense H from unit vectors, as NuTRL's ``matrixs`` builds it."""
        H = np.zeros((self.N, self.N))
        e = np.zeros(self.N)
        for i in range(self.N):
            e[i] = 1.0
            H[:, i] = self.apply(e)
            e[i] = 0.0
        return H


def sixj(a, b, c, d, e, f) -> float:
    return lib().refcpu_sixj(a, b, c, d, e, f)


def tfind(keys: bytes, key: bytes) -> int:
    n = len(keys) // 8
    kb = C.create_string_buffer(keys, len(keys) or 1)
    k = C.create_string_buffer(key, 8)
    return lib().refcpu_tfind(kb, n, k)
