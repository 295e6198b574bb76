"""ctypes binding of ``libnsx.so`` (the C ABI declared in ``include/nsx_hv.h``).

The library is built in-tree by ``nushellx_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for the H*v.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnsx.so")

NSX_FP64, NSX_FP32, NSX_BF16X3 = 0, 1, 2
NSX_HOST_ONLY = 0x100
NSX_TEMPLATE = 0x200
PRECISIONS = {"fp64": NSX_FP64, "fp32": NSX_FP32, "bf16x3": NSX_BF16X3}

# every symbol include/nsx_hv.h declares
EXPORTS = [
    "nsx_hv_create", "nsx_hv_destroy", "nsx_hv_dims", "nsx_hv_block", "nsx_hv_lcoef",
    "nsx_hv_work", "nsx_hv_apply_f32", "nsx_hv_apply_f64", "nsx_hv_apply_dev",
    "nsx_batch_create", "nsx_batch_destroy", "nsx_batch_dims", "nsx_batch_offset",
    "nsx_batch_work", "nsx_batch_apply_dev", "nsx_trl_solve", "nsx_dev_alloc", "nsx_dev_free",
    "nsx_dev_upload", "nsx_dev_download", "nsx_dev_sync", "nsx_batch_time", "nsx_device_info",
    "nsx_hv_autotune", "nsx_batch_autotune", "nsx_batch_set_config", "nsx_batch_path", "nsx_batch_trl_solve", "nsx_dev_copy", "nsx_batch_create_config", "nsx_hv_write_vectors",
    "nsx_hv_create_sep", "nsx_hv_write_xvc", "nsx_hv_write_trl", "nsx_nuvec", "nsx_hv_expect", "nsx_hv_expect_noalloc", "nsx_hv_xpct_files", "nsx_hv_trunc",
    "nsx_ens_layout", "nsx_ens_values", "nsx_ens_create", "nsx_ens_update", "nsx_ens_samples",
    "nsx_mlp_create", "nsx_mlp_destroy", "nsx_mlp_num_params", "nsx_mlp_get_params",
    "nsx_mlp_set_params", "nsx_mlp_grad", "nsx_mlp_train", "nsx_mlp_predict",
    "nsx_last_error",
]


class TrlOpts(C.Structure):
    _fields_ = [("nev", C.c_int32), ("max_basis", C.c_int32), ("keep", C.c_int32),
                ("max_iter", C.c_int32), ("tol", C.c_double), ("seed", C.c_uint64),
                ("check_every", C.c_int32), ("graph", C.c_int32), ("checkpoint", C.c_char_p),
                ("standard", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("fat_cap", C.c_int32), ("max_phases", C.c_int32), ("threads", C.c_int32),
                ("mode", C.c_int32), ("atomic", C.c_int32), ("ms", C.c_double),
                ("stacked", C.c_int32), ("stk_cap", C.c_int32), ("stk_work", C.c_int32),
                ("f_rounds", C.c_int32), ("deterministic", C.c_int32), ("big_rows1", C.c_int32),
                ("big_kchunk", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrlStats(C.Structure):
    _fields_ = [("matvecs", C.c_int32), ("restarts", C.c_int32), ("converged", C.c_int32),
                ("seconds", C.c_double), ("host_seconds", C.c_double)]


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m nushellx_amd.build` "
                           "(the H*v has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.nsx_hv_create.argtypes = [C.c_char_p] * 4 + [C.c_int, C.POINTER(vp)]
    L.nsx_hv_destroy.argtypes = [vp]
    L.nsx_hv_destroy.restype = None
    L.nsx_hv_dims.argtypes = [vp, C.POINTER(i64), C.POINTER(i32)]
    L.nsx_hv_block.argtypes = [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64), C.POINTER(i64)]
    L.nsx_hv_lcoef.argtypes = [vp, vp]
    L.nsx_hv_work.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(i64)]
    L.nsx_hv_apply_f32.argtypes = [vp, vp, vp]
    L.nsx_hv_apply_f64.argtypes = [vp, vp, vp]
    L.nsx_hv_apply_dev.argtypes = [vp, vp, vp, vp]
    L.nsx_batch_create.argtypes = [C.POINTER(vp), i32, C.c_int, C.POINTER(vp)]
    L.nsx_batch_destroy.argtypes = [vp]
    L.nsx_batch_destroy.restype = None
    L.nsx_batch_dims.argtypes = [vp, C.POINTER(i64)]
    L.nsx_batch_offset.argtypes = [vp, i32, C.POINTER(i64)]
    L.nsx_batch_work.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(i64)]
    L.nsx_batch_apply_dev.argtypes = [vp, vp, vp, vp]
    L.nsx_trl_solve.argtypes = [vp, C.POINTER(TrlOpts), vp, vp, vp, vp, C.POINTER(TrlStats)]
    L.nsx_dev_alloc.argtypes = [i64, C.POINTER(vp)]
    L.nsx_dev_free.argtypes = [vp]
    L.nsx_dev_upload.argtypes = [vp, vp, i64]
    L.nsx_dev_download.argtypes = [vp, vp, i64]
    L.nsx_dev_sync.argtypes = []
    L.nsx_batch_time.argtypes = [vp, vp, vp, i32, i32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
    L.nsx_device_info.argtypes = [C.c_char_p, i32, C.POINTER(i32)]
    L.nsx_hv_autotune.argtypes = [vp, i32, C.POINTER(TuneResult)]
    L.nsx_batch_autotune.argtypes = [vp, i32, C.POINTER(TuneResult)]
    L.nsx_batch_set_config.argtypes = [vp, C.POINTER(TuneResult)]
    L.nsx_batch_path.argtypes = [vp, C.POINTER(i32)]
    L.nsx_batch_create_config.argtypes = [C.POINTER(vp), i32, C.c_int, C.POINTER(TuneResult), C.POINTER(vp)]
    L.nsx_hv_write_vectors.argtypes = [vp, C.c_char_p, i32, vp, vp, i32, C.c_float]
    L.nsx_hv_trunc.argtypes = [vp, vp]
    L.nsx_hv_create_sep.argtypes = [C.c_char_p] * 5 + [C.c_int, C.POINTER(vp)]
    L.nsx_hv_write_xvc.argtypes = [vp, C.c_char_p, i32, vp, vp]
    L.nsx_hv_write_trl.argtypes = [vp, C.c_char_p, i32, vp, vp, i32, C.c_float]
    L.nsx_nuvec.argtypes = [vp, C.c_char_p, C.POINTER(i32)]
    L.nsx_hv_expect.argtypes = [vp, i32, vp, vp]
    L.nsx_hv_expect_noalloc.argtypes = [vp, i32, vp, vp]
    L.nsx_hv_xpct_files.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(i32)]
    L.nsx_dev_copy.argtypes = [vp, vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    L.nsx_batch_trl_solve.argtypes = [vp, C.POINTER(TrlOpts), vp, vp, vp, vp, vp, C.POINTER(TrlStats)]
    L.nsx_ens_layout.argtypes = [vp, C.POINTER(i64), C.POINTER(i32), vp, i32]
    L.nsx_ens_values.argtypes = [vp, vp]
    L.nsx_ens_create.argtypes = [C.POINTER(vp), i32, i32, i32, vp, C.POINTER(vp), C.c_int, C.POINTER(TuneResult),
                                 C.POINTER(vp)]
    L.nsx_ens_update.argtypes = [vp, vp]
    L.nsx_ens_samples.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
    L.nsx_mlp_create.argtypes = [i32, i32, i32, i32, C.c_uint64, C.POINTER(vp)]
    L.nsx_mlp_destroy.argtypes = [vp]
    L.nsx_mlp_destroy.restype = None
    L.nsx_mlp_num_params.argtypes = [vp, C.POINTER(i64)]
    L.nsx_mlp_get_params.argtypes = [vp, vp]
    L.nsx_mlp_set_params.argtypes = [vp, vp]
    L.nsx_mlp_grad.argtypes = [vp, vp, vp, i32, vp, C.POINTER(C.c_double)]
    L.nsx_mlp_train.argtypes = [vp, vp, vp, i32, i32, i32, C.c_float, C.c_uint64, vp]
    L.nsx_mlp_predict.argtypes = [vp, vp, i32, vp]
    L.nsx_last_error.restype = C.c_char_p
    _LIB = L
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().nsx_last_error().decode(errors="replace")
        raise NsxError(rc, f"{what} failed ({rc}): {msg}")


class NsxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class DeviceBuffer:
    """A device allocation owned through the library's own HIP runtime."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(lib().nsx_dev_alloc(int(nbytes), C.byref(p)), "nsx_dev_alloc")
        self.ptr = p
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a) -> None:
        import numpy as np
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().nsx_dev_upload(self.ptr, a.ctypes.data, a.nbytes), "nsx_dev_upload")

    def download(self, a) -> None:
        assert a.nbytes <= self.nbytes and a.flags["C_CONTIGUOUS"]
        check(lib().nsx_dev_download(a.ctypes.data, self.ptr, a.nbytes), "nsx_dev_download")

    def data_ptr(self) -> int:
        return int(self.ptr.value or 0)

    def free(self) -> None:
        if self.ptr:
            lib().nsx_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync() -> None:
    check(lib().nsx_dev_sync(), "nsx_dev_sync")


def stream_copy(nbytes: int, iters: int = 5) -> float:
    """Average ms of the library's streaming copy kernel over nbytes (both buffers on device)."""
    a, b = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    ms = C.c_double()
    try:
        check(lib().nsx_dev_copy(b.ptr, a.ptr, nbytes, iters, C.byref(ms)), "nsx_dev_copy")
    finally:
        a.free()
        b.free()
    return ms.value


def device_info():
    name = C.create_string_buffer(128)
    cus = C.c_int32()
    check(lib().nsx_device_info(name, 128, C.byref(cus)), "nsx_device_info")
    return name.value.decode(), int(cus.value)
