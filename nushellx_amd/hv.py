"""Host-side mirror of NuTRL's H*v interface over the C ABI.

NuTRL is launched as ``NuTRL <Nucleus+Jcode> <Nucleus>a <Nucleus>b <Nucleus><xtra>``
(``src/NuShellX/NuShellx.f90:1060-1063``) and, after its setup, calls
``multiplyx(); multiplya(); multiplyb()`` to turn ``coefR`` into ``coefL = H coefR``
(``src/NuTRL/NuCoreX.f90:317-319``).  ``Hamiltonian`` takes the same four names,
and ``multiply`` is that call sequence.  Errors surface as ``NsxError`` carrying
the negative code of ``include/nsx_hv.h`` (the reference stops with a message).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import (NSX_FP64, NSX_HOST_ONLY, NSX_TEMPLATE, PRECISIONS, DeviceBuffer, NsxError, TrlOpts, TrlStats, TuneResult, check,
                   device_sync, lib)


class Hamiltonian:
    def __init__(self, nus_prefix: str, opfile_a: str, opfile_b: str, intfile: str,
                 precision: str = "fp64", device: bool = True, dopfile: Optional[str] = None,
                 template: bool = False):
        """``device=False`` builds a host-only instance (no device memory), usable
        only as a member of a Batch.  ``dopfile`` takes the pn part from the
        separable operator ``<dopfile>.dop`` (program NuTRLZ / NuXpct).
        ``template``: a template of a linear ensemble (every .oph key kept in the
        term structure; see ``Batch.linear_ensemble``)."""
        L = lib()
        self.precision = precision
        self.device = device
        self.dopfile = dopfile
        h = C.c_void_p()
        flags = PRECISIONS[precision] | (0 if device else NSX_HOST_ONLY) | (NSX_TEMPLATE if template else 0)
        if dopfile is None:
            check(L.nsx_hv_create(nus_prefix.encode(), opfile_a.encode(), opfile_b.encode(),
                                  intfile.encode(), flags, C.byref(h)), "nsx_hv_create")
        else:
            check(L.nsx_hv_create_sep(nus_prefix.encode(), opfile_a.encode(), opfile_b.encode(),
                                      intfile.encode(), dopfile.encode(), flags, C.byref(h)),
                  "nsx_hv_create_sep")
        self._h = h
        n, b = C.c_int64(), C.c_int32()
        check(L.nsx_hv_dims(h, C.byref(n), C.byref(b)), "nsx_hv_dims")
        self.N, self.basd = int(n.value), int(b.value)

    def ens_layout(self):
        """Value layout of a template: (number of .oph records, [(side, 2J, d), ...])
        of its H blocks (side 0 = a, 1 = b)."""
        no, nh = C.c_int64(), C.c_int32()
        check(lib().nsx_ens_layout(self._h, C.byref(no), C.byref(nh), None, 0), "nsx_ens_layout")
        desc = np.zeros(3 * max(nh.value, 1), dtype=np.int32)
        check(lib().nsx_ens_layout(self._h, C.byref(no), C.byref(nh), desc.ctypes.data, nh.value), "nsx_ens_layout")
        return int(no.value), [tuple(int(x) for x in desc[3 * i:3 * i + 3]) for i in range(nh.value)]

    def ens_values(self) -> np.ndarray:
        """The template's own file values in its ensemble layout."""
        no, hl = self.ens_layout()
        out = np.zeros(no + sum(d * d for _, _, d in hl))
        check(lib().nsx_ens_values(self._h, out.ctypes.data), "nsx_ens_values")
        return out

    @classmethod
    def from_args(cls, args: Sequence[str], precision: str = "fp64") -> "Hamiltonian":
        return cls(*args, precision=precision)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nsx_hv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- basis ---------------------------------------------------------------
    def blocks(self) -> List[Tuple[int, int, int, int]]:
        out = []
        a, b, p, d = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int64()
        for i in range(self.basd):
            check(lib().nsx_hv_block(self._h, i, C.byref(a), C.byref(b), C.byref(p), C.byref(d)),
                  "nsx_hv_block")
            out.append((a.value, b.value, p.value, d.value))
        return out

    def lcoef(self) -> np.ndarray:
        m = np.zeros(max(self.N, 1), dtype=np.uint8)
        check(lib().nsx_hv_lcoef(self._h, m.ctypes.data), "nsx_hv_lcoef")
        return m[:self.N].astype(bool)

    def write_vectors(self, prefix: str, evals, vecs, no_level: int = 5, tt: float = 0.0) -> None:
        """<prefix>.lnz / <prefix>.abz in the layout NuVec reads (vecs: nvec x N)."""
        vecs = np.ascontiguousarray(np.atleast_2d(vecs), dtype=np.float64)
        evals = np.ascontiguousarray(evals, dtype=np.float64)
        assert vecs.shape == (len(evals), self.N)
        check(lib().nsx_hv_write_vectors(self._h, prefix.encode(), len(evals), evals.ctypes.data,
                                         vecs.ctypes.data, no_level, C.c_float(tt)), "nsx_hv_write_vectors")

    def write_trl(self, prefix: str, evals, vecs, no_level: int = 5, tt: float = 0.0) -> None:
        """<prefix>.trl / <prefix>.abz in the thick-restart layout NuVec reads (vecs: nvec x N)."""
        vecs = np.ascontiguousarray(np.atleast_2d(vecs), dtype=np.float64)
        evals = np.ascontiguousarray(evals, dtype=np.float64)
        assert vecs.shape == (len(evals), self.N)
        check(lib().nsx_hv_write_trl(self._h, prefix.encode(), len(evals), evals.ctypes.data,
                                     vecs.ctypes.data, no_level, C.c_float(tt)), "nsx_hv_write_trl")

    def nuvec(self, prefix: str) -> int:
        """NuVec's step: <prefix>.abz + .lnz/.trl in, .xvc/.xva/.lev/.lp out; returns the state count."""
        n = C.c_int32()
        check(lib().nsx_nuvec(self._h, prefix.encode(), C.byref(n)), "nsx_nuvec")
        return int(n.value)

    def trunc(self) -> Tuple[bool, np.ndarray]:
        """(truncated, mask): the states a solve spans (the N-hbar-omega window
        mask on a truncated basis, else the parity mask)."""
        m = np.zeros(self.N, dtype=np.uint8)
        rc = lib().nsx_hv_trunc(self._h, m.ctypes.data)
        if rc < 0:
            check(rc, "nsx_hv_trunc")
        return bool(rc), m.astype(bool)

    def write_xvc(self, prefix: str, evals, vecs) -> None:
        """<prefix>.xvc, NuVec's state file (vecs: nvec x N)."""
        vecs = np.ascontiguousarray(np.atleast_2d(vecs), dtype=np.float64)
        evals = np.ascontiguousarray(evals, dtype=np.float64)
        assert vecs.shape == (len(evals), self.N)
        check(lib().nsx_hv_write_xvc(self._h, prefix.encode(), len(evals), evals.ctypes.data,
                                     vecs.ctypes.data), "nsx_hv_write_xvc")

    def expect(self, vecs) -> np.ndarray:
        """M[i, j] = (H v_i) . v_j for the rows v_i of ``vecs`` (NuXpct's matrixs)."""
        vecs = np.ascontiguousarray(np.atleast_2d(vecs), dtype=np.float64)
        n = vecs.shape[0]
        assert vecs.shape[1] == self.N
        M = np.empty((n, n), dtype=np.float64)  # C order: M[j, i] holds the (i, j) element
        check(lib().nsx_hv_expect(self._h, n, vecs.ctypes.data, M.ctypes.data), "nsx_hv_expect")
        return M.T.copy()

    def xpct_files(self, prefix: str, operat: str) -> int:
        """NuXpct's file step: <prefix>.xvc in, <prefix><operat>.ovl/.ovx out; returns the level count."""
        n = C.c_int32()
        check(lib().nsx_hv_xpct_files(self._h, prefix.encode(), operat.encode(), C.byref(n)),
              "nsx_hv_xpct_files")
        return int(n.value)

    def work(self) -> Tuple[float, float, int]:
        """(flops, like-particle flops, FP64 bytes) of one H*v (algorithmic counts)."""
        f, fl, b = C.c_double(), C.c_double(), C.c_int64()
        check(lib().nsx_hv_work(self._h, C.byref(f), C.byref(fl), C.byref(b)), "nsx_hv_work")
        return f.value, fl.value, int(b.value)

    # ---- H*v -----------------------------------------------------------------
    def multiply(self, v: np.ndarray) -> np.ndarray:
        """coefL = H coefR on host vectors (float32 in -> float32 out, else float64)."""
        if v.shape != (self.N,):
            raise ValueError(f"vector of length {self.N} expected, got {v.shape}")
        if v.dtype == np.float32:
            v = np.ascontiguousarray(v)
            out = np.empty_like(v)
            check(lib().nsx_hv_apply_f32(self._h, v.ctypes.data, out.ctypes.data), "nsx_hv_apply_f32")
            return out
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty_like(v)
        check(lib().nsx_hv_apply_f64(self._h, v.ctypes.data, out.ctypes.data), "nsx_hv_apply_f64")
        return out

    def multiply_dev(self, d_in, d_out, stream=None) -> None:
        """H*v on device buffers (``DeviceBuffer`` or raw pointers) on a HIP stream."""
        pin = d_in.data_ptr() if hasattr(d_in, "data_ptr") else int(d_in)
        pout = d_out.data_ptr() if hasattr(d_out, "data_ptr") else int(d_out)
        s = None if stream is None else (stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        check(lib().nsx_hv_apply_dev(self._h, C.c_void_p(pin), C.c_void_p(pout), C.c_void_p(s)),
              "nsx_hv_apply_dev")

    def autotune(self, iters: int = 10) -> dict:
        """Pick the staged-kernel shape by timing the candidates on this instance."""
        r = TuneResult()
        check(lib().nsx_hv_autotune(self._h, iters, C.byref(r)), "nsx_hv_autotune")
        return r.as_dict()

    # ---- Lanczos ----------------------------------------------------------------
    def lanczos(self, nev: int = 5, max_basis: int = 50, keep: int = 10, tol: float = 1e-6,
                max_iter: int = 5000, seed: int = 5, start: Optional[np.ndarray] = None,
                vectors: bool = False, check_every: int = 8, graph: bool = True,
                checkpoint: Optional[str] = None, standard: bool = False):
        """Thick-restart Lanczos for the lowest ``nev`` states (NuTRL's Lnz_OP);
        ``standard=True``: standard Lanczos without restarts (Lnzs), converged
        when the lowest nev+1 Ritz values move by <= tol between tests.

        Convergence is tested every ``check_every`` steps; with ``graph`` each
        block of steps is captured once as a hipGraph and replayed.  With
        ``checkpoint`` the restart state is saved to that file at every thick
        restart, and a matching file found at the start is resumed."""
        opts = TrlOpts(nev, max_basis, keep, max_iter, tol, seed, check_every, int(graph),
                       checkpoint.encode() if checkpoint else None, int(standard))
        ev = np.zeros(nev)
        res = np.zeros(nev)
        vec = np.zeros((nev, self.N)) if vectors else None  # column-major N x nev == row-major nev x N
        st = TrlStats()
        s = None
        if start is not None:
            start = np.ascontiguousarray(start, dtype=np.float64)
            s = start.ctypes.data
        check(lib().nsx_trl_solve(self._h, C.byref(opts), s, ev.ctypes.data,
                                  vec.ctypes.data if vectors else None, res.ctypes.data, C.byref(st)),
              "nsx_trl_solve")
        out = {"evals": ev, "residuals": res, "matvecs": st.matvecs, "restarts": st.restarts,
               "converged": st.converged, "seconds": st.seconds,
               "host_seconds": st.host_seconds}
        if vectors:
            out["evecs"] = vec
        return out


class Batch:
    """Many instances in one launch pair (vectors concatenated in instance order)."""

    def __init__(self, hams: Sequence[Hamiltonian], precision: str = "fp64", config: Optional[dict] = None):
        """``config``: a result of ``autotune`` (plan shape and kernel choice) to
        plan with from the start instead of the defaults."""
        arr = (C.c_void_p * len(hams))(*[h.handle for h in hams])
        b = C.c_void_p()
        if config is None:
            check(lib().nsx_batch_create(arr, len(hams), PRECISIONS[precision], C.byref(b)), "nsx_batch_create")
        else:
            cfg = TuneResult(**{k: config.get(k, 0) for k, _ in TuneResult._fields_})
            check(lib().nsx_batch_create_config(arr, len(hams), PRECISIONS[precision], C.byref(cfg), C.byref(b)),
                  "nsx_batch_create_config")
        self._b = b
        self._hams = list(hams)
        self.samples = 0
        self._init_dims()

    @classmethod
    def linear_ensemble(cls, templates: Sequence[Hamiltonian], theta: np.ndarray, bases: Sequence[np.ndarray],
                        precision: str = "fp64", config: Optional[dict] = None) -> "Batch":
        """S interaction samples (theta: S x P) of every template in one batch, the
        operators of each sample made on the device from the templates' bases
        (basis t: P x K_t, the template's values for each unit parameter vector,
        in the layout of ``Hamiltonian.ens_layout``).  Instances are sample-major:
        instance s * len(templates) + t is template t with sample s."""
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        S, P = theta.shape
        bases = [np.ascontiguousarray(B, dtype=np.float64) for B in bases]
        assert len(bases) == len(templates) and all(B.shape[0] == P for B in bases)
        arr = (C.c_void_p * len(templates))(*[h.handle for h in templates])
        bp = (C.c_void_p * len(bases))(*[B.ctypes.data for B in bases])
        cfg = None
        if config is not None:
            cfg = TuneResult(**{k: config.get(k, 0) for k, _ in TuneResult._fields_})
        b = C.c_void_p()
        check(lib().nsx_ens_create(arr, len(templates), S, P, theta.ctypes.data, bp, PRECISIONS[precision],
                                   C.byref(cfg) if cfg is not None else None, C.byref(b)), "nsx_ens_create")
        self = cls.__new__(cls)
        self._b = b
        self._hams = [h for _ in range(S) for h in templates]
        self._bases = bases
        self.samples = S
        self._init_dims()
        return self

    def ens_update(self, theta: np.ndarray) -> None:
        """New parameters for the same samples; operators rebuilt on the device."""
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        assert theta.shape[0] == self.samples
        check(lib().nsx_ens_update(self._b, theta.ctypes.data), "nsx_ens_update")

    def _init_dims(self) -> None:
        b = self._b
        n = C.c_int64()
        check(lib().nsx_batch_dims(b, C.byref(n)), "nsx_batch_dims")
        self.N = int(n.value)
        self.offsets = []
        for i in range(len(self._hams)):
            o = C.c_int64()
            check(lib().nsx_batch_offset(b, i, C.byref(o)), "nsx_batch_offset")
            self.offsets.append(int(o.value))

    def work(self) -> Tuple[float, float, int]:
        f, fl, by = C.c_double(), C.c_double(), C.c_int64()
        check(lib().nsx_batch_work(self._b, C.byref(f), C.byref(fl), C.byref(by)), "nsx_batch_work")
        return f.value, fl.value, int(by.value)

    def multiply_dev(self, d_in, d_out, stream=None) -> None:
        pin = d_in.data_ptr() if hasattr(d_in, "data_ptr") else int(d_in)
        pout = d_out.data_ptr() if hasattr(d_out, "data_ptr") else int(d_out)
        s = None if stream is None else (stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        check(lib().nsx_batch_apply_dev(self._b, C.c_void_p(pin), C.c_void_p(pout), C.c_void_p(s)),
              "nsx_batch_apply_dev")

    def autotune(self, iters: int = 10) -> dict:
        r = TuneResult()
        check(lib().nsx_batch_autotune(self._b, iters, C.byref(r)), "nsx_batch_autotune")
        return r.as_dict()

    def path(self) -> str:
        """The small-block pn path of the plan: "staged", "stacked", "fused",
        "stacked-mfma" (the stacked stages on the matrix cores) or "fused-mfma"
        (both stages in one matrix-core kernel; " +like" when it also carries
        the like-particle part), with " deterministic" for the plans of the
        deterministic mode."""
        v = C.c_int32()
        check(lib().nsx_batch_path(self._b, C.byref(v)), "nsx_batch_path")
        if v.value & 8:
            return "fused-mfma" + (" +like" if v.value & 16 else "") + (" deterministic" if v.value & 4 else "")
        return ("staged", "stacked", "fused", "stacked-mfma")[v.value & 3] + (" deterministic" if v.value & 4 else "")

    def set_config(self, cfg: dict) -> None:
        r = TuneResult(cfg["fat_cap"], cfg["max_phases"], cfg["threads"], cfg["mode"],
                       cfg.get("atomic", 0), 0.0, cfg.get("stacked", 0), cfg.get("stk_cap", 0),
                       cfg.get("stk_work", 0), cfg.get("f_rounds", 0), cfg.get("deterministic", 0),
                       cfg.get("big_rows1", 0), cfg.get("big_kchunk", 0))
        check(lib().nsx_batch_set_config(self._b, C.byref(r)), "nsx_batch_set_config")

    def lanczos(self, nev: int = 5, max_basis: int = 50, keep: int = 10, tol: float = 1e-6,
                max_iter: int = 5000, seed: int = 5, start: Optional[np.ndarray] = None,
                vectors: bool = False, check_every: int = 8, graph: bool = True,
                checkpoint: Optional[str] = None, standard: bool = False):
        """Lockstep thick-restart Lanczos over every instance of the batch.

        ``evecs`` (with ``vectors``) is nev x N_total: row q holds every
        instance's q-th Ritz vector at that instance's offset."""
        n = len(self._hams)
        opts = TrlOpts(nev, max_basis, keep, max_iter, tol, seed, check_every, int(graph),
                       checkpoint.encode() if checkpoint else None, int(standard))
        vec = np.zeros((nev, self.N)) if vectors else None
        ev = np.zeros((n, nev))
        res = np.zeros((n, nev))
        conv = np.zeros(n, dtype=np.int32)
        st = TrlStats()
        s = None
        if start is not None:
            start = np.ascontiguousarray(start, dtype=np.float64)
            assert start.shape == (self.N,)
            s = start.ctypes.data
        check(lib().nsx_batch_trl_solve(self._b, C.byref(opts), s, ev.ctypes.data,
                                        vec.ctypes.data if vectors else None, res.ctypes.data,
                                        conv.ctypes.data, C.byref(st)), "nsx_batch_trl_solve")
        out = {"evals": ev, "residuals": res, "converged_at": conv, "matvecs": st.matvecs,
               "restarts": st.restarts, "converged": st.converged, "seconds": st.seconds,
               "host_seconds": st.host_seconds}
        if vectors:
            out["evecs"] = vec
        return out

    def write_vectors(self, prefixes: Sequence[str], result: dict, no_level: int = 5) -> None:
        """Per instance, <prefix>.lnz/.abz (NuVec layout) from ``lanczos(vectors=True)``."""
        for k, (h, prefix) in enumerate(zip(self._hams, prefixes)):
            o = self.offsets[k]
            h.write_vectors(prefix, result["evals"][k], result["evecs"][:, o:o + h.N], no_level=no_level)

    def time(self, d_in, d_out, warmup: int = 10, iters: int = 100):
        """HIP-event timing (ms per batched H*v, pn kernel, reduce kernel)."""
        hv, pn, red = C.c_double(), C.c_double(), C.c_double()
        check(lib().nsx_batch_time(self._b, C.c_void_p(d_in.data_ptr()), C.c_void_p(d_out.data_ptr()),
                                   warmup, iters, C.byref(hv), C.byref(pn), C.byref(red)),
              "nsx_batch_time")
        return hv.value, pn.value, red.value

    def close(self):
        if getattr(self, "_b", None):
            lib().nsx_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Hamiltonian", "Batch", "DeviceBuffer", "device_sync", "NsxError", "NSX_FP64"]
