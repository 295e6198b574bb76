"""The Fortran binding (include/nsx_hv.f90) compiles with the image's flang and
drives the library from Fortran: a create call on missing files returns NSX_EIO
with the library's message, before any device call (no GPU needed)."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or shutil.which("flang")
LIB = os.path.join(ROOT, "nushellx_amd", "libnsx.so")

PROGRAM = """
program nsx_error_path
  use iso_c_binding
  use nsx_hv_api
  implicit none
  type(c_ptr) :: h
  integer(c_int) :: rc
  character(len=40) :: nucleus = '/nonexistent/ne200   '
  type(nsx_trl_opts) :: opt
  rc = nsx_hv_create(nsx_cstr(nucleus), nsx_cstr('/nonexistent/ne20a'), nsx_cstr('/nonexistent/ne20b'), &
                     nsx_cstr('/nonexistent/ne20x'), NSX_FP32, h)
  print '(i0)', rc
  print '(a)', nsx_error_message()
  print '(i0,1x,i0,1x,i0)', opt%nev, opt%max_basis, opt%keep
  rc = nsx_hv_create_sep(nsx_cstr(nucleus), nsx_cstr('/nonexistent/ne20a'), nsx_cstr('/nonexistent/ne20b'), &
                         nsx_cstr('/nonexistent/ne20x'), nsx_cstr('/nonexistent/dz'), NSX_FP32, h)
  print '(i0)', rc
end program nsx_error_path
"""

# NuTRLZ's setup through the Fortran binding: a host-only separable handle
# (no device memory) on real files, and its basis dimension.
PROGRAM_SEP = """
program nsx_sep_setup
  use iso_c_binding
  use nsx_hv_api
  implicit none
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer(c_int64_t) :: n
  integer(c_int32_t) :: basd
  character(len=512) :: a(5)
  integer :: i
  do i = 1, 5
    call get_command_argument(i, a(i))
  end do
  rc = nsx_hv_create_sep(nsx_cstr(a(1)), nsx_cstr(a(2)), nsx_cstr(a(3)), nsx_cstr(a(4)), nsx_cstr(a(5)), &
                         NSX_FP64 + NSX_HOST_ONLY, h)
  print '(i0)', rc
  if (rc /= NSX_OK) print '(a)', nsx_error_message()
  rc = nsx_hv_dims(h, n, basd)
  print '(i0,1x,i0)', n, basd
end program nsx_sep_setup
"""


@pytest.mark.skipif(FC is None or not os.path.exists(LIB), reason="no Fortran compiler or library not built")
def test_fortran_binding_error_path():
    d = tempfile.mkdtemp(prefix="nsx_f90_")
    src = os.path.join(d, "prog.f90")
    with open(src, "w") as f:
        f.write(PROGRAM)
    exe = os.path.join(d, "prog")
    libdir = os.path.dirname(LIB)
    subprocess.run([FC, "-O1", "-J", d, os.path.join(ROOT, "include", "nsx_hv.f90"), src, "-o", exe,
                    f"-L{libdir}", "-lnsx", f"-Wl,-rpath,{libdir}"], check=True, cwd=d,
                   capture_output=True, timeout=300)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, cwd=d).stdout.split("\n")
    assert out[0].strip() == "-1", out
    assert "cannot open" in out[1] and "/nonexistent/ne200.nus" in out[1], out
    assert out[2].split() == ["5", "50", "10"]
    assert out[3].strip() == "-1", out


@pytest.mark.skipif(FC is None or not os.path.exists(LIB), reason="no Fortran compiler or library not built")
def test_fortran_separable_setup(models, dops):
    from oracle.refcpu import RefCPU
    d = tempfile.mkdtemp(prefix="nsx_f90_")
    src = os.path.join(d, "prog.f90")
    with open(src, "w") as f:
        f.write(PROGRAM_SEP)
    exe = os.path.join(d, "prog")
    libdir = os.path.dirname(LIB)
    subprocess.run([FC, "-O1", "-J", d, os.path.join(ROOT, "include", "nsx_hv.f90"), src, "-o", exe,
                    f"-L{libdir}", "-lnsx", f"-Wl,-rpath,{libdir}"], check=True, cwd=d,
                   capture_output=True, timeout=300)
    args = models["sd23"].args(3)
    out = subprocess.run([exe, *args, dops["sd23"]["cut"]], capture_output=True, text=True, timeout=120,
                         cwd=d).stdout.split("\n")
    assert out[0].strip() == "0", out
    ref = RefCPU(*args)
    assert out[1].split() == [str(ref.N), str(ref.basd)]
