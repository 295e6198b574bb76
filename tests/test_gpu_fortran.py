"""The REAL*4 drop-in driven from Fortran on the GPU, as NuTRL's Lanczos would
call it: an amdflang program creates a handle on the operator files through
include/nsx_hv.f90, calls nsx_hv_apply_f32 with host coefR -> coefL (the place
of `call multiplyx(); call multiplya(); call multiplyb()`,
src/NuTRL/NuCoreX.f90:317-319), writes coefL, and times repeated calls.  The
result equals the FP64 oracle's H*v of the same REAL*4 input to REAL*4
rounding."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from oracle.refcpu import RefCPU

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or shutil.which("flang")
LIB = os.path.join(ROOT, "nushellx_amd", "libnsx.so")

PROGRAM = """
program nsx_nutrl_hv
  use iso_c_binding
  use nsx_hv_api
  implicit none
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer(c_int64_t) :: n
  integer(c_int32_t) :: basd
  character(len=512) :: a(5)
  real(c_float), allocatable :: coefR(:), coefL(:)
  integer :: i, it
  integer, parameter :: niter = 20
  integer(8) :: t0, t1, rate
  do i = 1, 5
    call get_command_argument(i, a(i))
  end do
  rc = nsx_hv_create(nsx_cstr(a(1)), nsx_cstr(a(2)), nsx_cstr(a(3)), nsx_cstr(a(4)), NSX_FP64, h)
  if (rc /= NSX_OK) then
    print '(a)', nsx_error_message()
    stop 1
  end if
  rc = nsx_hv_dims(h, n, basd)
  allocate(coefR(n), coefL(n))
  do i = 1, int(n)
    coefR(i) = real(sin(0.37d0 * i), c_float)
  end do
  ! in place of: call multiplyx(); call multiplya(); call multiplyb()
  rc = nsx_hv_apply_f32(h, coefR, coefL)
  if (rc /= NSX_OK) then
    print '(a)', nsx_error_message()
    stop 2
  end if
  open(unit=11, file=trim(a(5)), form='unformatted', access='stream', status='replace')
  write(11) coefL
  close(11)
  call system_clock(t0, rate)
  do it = 1, niter
    rc = nsx_hv_apply_f32(h, coefR, coefL)
  end do
  call system_clock(t1)
  print '(i0,1x,i0,1x,f14.3)', n, basd, 1.0d6 * real(t1 - t0, 8) / real(rate, 8) / niter
  call nsx_hv_destroy(h)
end program nsx_nutrl_hv
"""


def _build(d):
    src = os.path.join(d, "prog.f90")
    with open(src, "w") as f:
        f.write(PROGRAM)
    exe = os.path.join(d, "prog")
    libdir = os.path.dirname(LIB)
    subprocess.run([FC, "-O2", "-J", d, os.path.join(ROOT, "include", "nsx_hv.f90"), src, "-o", exe,
                    f"-L{libdir}", "-lnsx", f"-Wl,-rpath,{libdir}"], check=True, cwd=d, capture_output=True,
                   timeout=300)
    return exe


@pytest.mark.skipif(FC is None, reason="no Fortran compiler")
def test_fortran_host_drop_in(models, pf44):
    d = tempfile.mkdtemp(prefix="nsx_f90gpu_")
    exe = _build(d)
    for name, args in (("sd44 J=4", models["sd44"].args(8)), ("sd66 J=2", models["sd66"].args(4)),
                       ("pf44 J=0", pf44.args(0))):
        out_file = os.path.join(d, "coefL.bin")
        print(f"case fortran {name}", flush=True)
        r = subprocess.run([exe, *args, out_file], capture_output=True, text=True, timeout=300, cwd=d)
        assert r.returncode == 0, r.stdout + r.stderr
        n, basd, us = r.stdout.split()
        ref = RefCPU(*args, fp64=True)
        assert int(n) == ref.N and int(basd) == ref.basd
        coefR = np.sin(0.37 * np.arange(1, ref.N + 1)).astype(np.float32)
        w = np.fromfile(out_file, dtype=np.float32)
        w_ref = ref.apply(coefR.astype(np.float64))
        err = np.abs(w - w_ref).max() / np.abs(w_ref).max()
        print(f"{name}: N={n}, {float(us):.1f} us per REAL*4 H*v from Fortran (host vectors), rel err {err:.1e}",
              flush=True)
        assert err < 1e-6
