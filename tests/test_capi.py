"""The C-ABI library loads and exports every function include/nsx_hv.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from nushellx_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "nsx_hv.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nsx_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name


def test_error_path_without_device():
    """Bad file names fail with NSX_EIO before any device call."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.nsx_hv_create(b"/nonexistent/x", b"/nonexistent/a", b"/nonexistent/b", b"/nonexistent/i",
                         0, ctypes.byref(h))
    assert rc == -1
    assert b"cannot open" in L.nsx_last_error()
