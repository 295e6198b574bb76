! nsx_hv.f90 -- Fortran 2003 binding (iso_c_binding) of include/nsx_hv.h.
!
! For the NuTRL host: create one handle per nucleus and J after the argument
! parsing of Lnz_TRD (src/NuTRL/NuCoreX.f90:1634-1670), replace the three calls
!   call multiplyx(); call multiplya(); call multiplyb()
! (src/NuTRL/NuCoreX.f90:317-319, 340-342) by one nsx_hv_apply_f32 call with
! coefR as input and coefL as output, and map a nonzero return code to `stop`.
! File names are passed through nsx_cstr, which trims the blank padding of
! Fortran character variables and appends the C terminator.
module nsx_hv_api
  use iso_c_binding
  implicit none
  private
  public :: nsx_hv_create, nsx_hv_create_sep, nsx_hv_destroy, nsx_hv_dims, nsx_hv_lcoef, nsx_hv_trunc, &
            nsx_hv_apply_f32, nsx_hv_apply_f64, nsx_trl_solve, nsx_hv_write_vectors, &
            nsx_hv_write_xvc, nsx_hv_expect, nsx_hv_xpct_files, nsx_last_error, nsx_cstr, &
            nsx_error_message, &
            nsx_trl_opts, nsx_trl_stats

  integer(c_int), parameter, public :: NSX_OK = 0, NSX_EIO = -1, NSX_EFORMAT = -2, NSX_ETAG = -3, &
                                       NSX_ECAP = -4, NSX_EGPU = -5, NSX_EARG = -6
  integer(c_int), parameter, public :: NSX_FP64 = 0, NSX_FP32 = 1, NSX_BF16X3 = 2
  integer(c_int), parameter, public :: NSX_HOST_ONLY = 256   ! or-ed into the precision
  integer(c_int), parameter, public :: NSX_TEMPLATE = 512    ! or-ed: a linear-ensemble template

  ! C: nsx_trl_opts (seed is uint64_t in C; any 64-bit pattern is accepted)
  type, bind(C) :: nsx_trl_opts
    integer(c_int32_t) :: nev = 5, max_basis = 50, keep = 10, max_iter = 5000
    real(c_double) :: tol = 1.0d-6
    integer(c_int64_t) :: seed = 5
    integer(c_int32_t) :: check_every = 8, graph = 1
    type(c_ptr) :: checkpoint = c_null_ptr   ! c_loc of a NUL-terminated file name, or null
    integer(c_int32_t) :: standard = 0       ! 1: standard Lanczos (Lnzs)
  end type nsx_trl_opts

  type, bind(C) :: nsx_trl_stats
    integer(c_int32_t) :: matvecs, restarts, converged
    real(c_double) :: seconds, host_seconds
  end type nsx_trl_stats

  interface
    integer(c_int) function nsx_hv_create(nus, opfile_a, opfile_b, intfile, precision, h) &
        bind(C, name='nsx_hv_create')
      import :: c_int, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: nus, opfile_a, opfile_b, intfile
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: h
    end function nsx_hv_create

    ! NuTRLZ / NuXpct: the pn part from the separable operator <dopfile>.dop
    integer(c_int) function nsx_hv_create_sep(nus, opfile_a, opfile_b, intfile, dopfile, precision, h) &
        bind(C, name='nsx_hv_create_sep')
      import :: c_int, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: nus, opfile_a, opfile_b, intfile, dopfile
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: h
    end function nsx_hv_create_sep

    subroutine nsx_hv_destroy(h) bind(C, name='nsx_hv_destroy')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine nsx_hv_destroy

    integer(c_int) function nsx_hv_dims(h, n, basd) bind(C, name='nsx_hv_dims')
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: n
      integer(c_int32_t), intent(out) :: basd
    end function nsx_hv_dims

    integer(c_int) function nsx_hv_lcoef(h, mask) bind(C, name='nsx_hv_lcoef')
      import :: c_int, c_ptr, c_int8_t
      type(c_ptr), value :: h
      integer(c_int8_t), dimension(*), intent(out) :: mask
    end function nsx_hv_lcoef

    ! states a solve spans: the N-hbar-omega window mask (returns 1) or lcoef (returns 0)
    integer(c_int) function nsx_hv_trunc(h, mask) bind(C, name='nsx_hv_trunc')
      import :: c_int, c_ptr, c_int8_t
      type(c_ptr), value :: h
      integer(c_int8_t), dimension(*), intent(out) :: mask
    end function nsx_hv_trunc

    ! coefL = H coefR (REAL*4 vectors of no_state elements; coefL is overwritten)
    integer(c_int) function nsx_hv_apply_f32(h, v_in, v_out) bind(C, name='nsx_hv_apply_f32')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), dimension(*), intent(in) :: v_in
      real(c_float), dimension(*), intent(out) :: v_out
    end function nsx_hv_apply_f32

    integer(c_int) function nsx_hv_apply_f64(h, v_in, v_out) bind(C, name='nsx_hv_apply_f64')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), dimension(*), intent(in) :: v_in
      real(c_double), dimension(*), intent(out) :: v_out
    end function nsx_hv_apply_f64

    ! start, evecs, residuals may be c_null_ptr (or c_loc of host arrays)
    integer(c_int) function nsx_trl_solve(h, opt, start, evals, evecs, residuals, stats) &
        bind(C, name='nsx_trl_solve')
      import :: c_int, c_ptr, c_double, nsx_trl_opts, nsx_trl_stats
      type(c_ptr), value :: h
      type(nsx_trl_opts), intent(in) :: opt
      type(c_ptr), value :: start
      real(c_double), dimension(*), intent(out) :: evals
      type(c_ptr), value :: evecs, residuals
      type(nsx_trl_stats), intent(out) :: stats
    end function nsx_trl_solve

    ! <prefix>.lnz / <prefix>.abz in the layout NuVec reads (vecs: N x nvec, column-major)
    integer(c_int) function nsx_hv_write_vectors(h, prefix, nvec, evals, vecs, no_level, tt) &
        bind(C, name='nsx_hv_write_vectors')
      import :: c_int, c_ptr, c_char, c_int32_t, c_double, c_float
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*), intent(in) :: prefix
      integer(c_int32_t), value :: nvec, no_level
      real(c_double), dimension(*), intent(in) :: evals, vecs
      real(c_float), value :: tt
    end function nsx_hv_write_vectors

    ! <prefix>.xvc as NuVec writes it (vecs: N x nvec, column-major)
    integer(c_int) function nsx_hv_write_xvc(h, prefix, nvec, evals, vecs) bind(C, name='nsx_hv_write_xvc')
      import :: c_int, c_ptr, c_char, c_int32_t, c_double
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*), intent(in) :: prefix
      integer(c_int32_t), value :: nvec
      real(c_double), dimension(*), intent(in) :: evals, vecs
    end function nsx_hv_write_xvc

    ! M(i,j) = (H v_i) . v_j (NuXpct's matrixs), M: nvec x nvec
    integer(c_int) function nsx_hv_expect(h, nvec, vecs, m) bind(C, name='nsx_hv_expect')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nvec
      real(c_double), dimension(*), intent(in) :: vecs
      real(c_double), dimension(*), intent(out) :: m
    end function nsx_hv_expect

    ! NuXpct's file step: <prefix>.xvc in, <prefix><operat>.ovl/.ovx out
    integer(c_int) function nsx_hv_xpct_files(h, prefix, operat, nlev) bind(C, name='nsx_hv_xpct_files')
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*), intent(in) :: prefix, operat
      integer(c_int32_t), intent(out) :: nlev
    end function nsx_hv_xpct_files

    ! <prefix>.trl / four-record <prefix>.abz, the thick-restart layout (vecs: N x nvec)
    integer(c_int) function nsx_hv_write_trl(h, prefix, nvec, evals, vecs, no_level, tt) &
        bind(C, name='nsx_hv_write_trl')
      import :: c_int, c_ptr, c_char, c_int32_t, c_double, c_float
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*), intent(in) :: prefix
      integer(c_int32_t), value :: nvec, no_level
      real(c_double), dimension(*), intent(in) :: evals, vecs
      real(c_float), value :: tt
    end function nsx_hv_write_trl

    ! NuVec's step: <prefix>.abz with .lnz or .trl in, .xvc/.xva/.lev/.lp out
    integer(c_int) function nsx_nuvec(h, prefix, nlev) bind(C, name='nsx_nuvec')
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*), intent(in) :: prefix
      integer(c_int32_t), intent(out) :: nlev
    end function nsx_nuvec

    ! batches: hs is an array of handles (type(c_ptr)), instances in that order
    integer(c_int) function nsx_batch_create(hs, n, precision, b) bind(C, name='nsx_batch_create')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), dimension(*), intent(in) :: hs
      integer(c_int32_t), value :: n
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: b
    end function nsx_batch_create

    subroutine nsx_batch_destroy(b) bind(C, name='nsx_batch_destroy')
      import :: c_ptr
      type(c_ptr), value :: b
    end subroutine nsx_batch_destroy

    ! evals, residuals: n*nev; converged_at: n; start, evecs may be c_null_ptr
    integer(c_int) function nsx_batch_trl_solve(b, opt, start, evals, evecs, residuals, converged_at, stats) &
        bind(C, name='nsx_batch_trl_solve')
      import :: c_int, c_ptr, c_double, c_int32_t, nsx_trl_opts, nsx_trl_stats
      type(c_ptr), value :: b
      type(nsx_trl_opts), intent(in) :: opt
      type(c_ptr), value :: start
      real(c_double), dimension(*), intent(out) :: evals
      type(c_ptr), value :: evecs
      real(c_double), dimension(*), intent(out) :: residuals
      integer(c_int32_t), dimension(*), intent(out) :: converged_at
      type(nsx_trl_stats), intent(out) :: stats
    end function nsx_batch_trl_solve

    ! linear ensembles: n_oph .oph records and n_h H blocks (side, 2J, d) of a template
    integer(c_int) function nsx_ens_layout(h, n_oph, n_h, h_desc, cap) bind(C, name='nsx_ens_layout')
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: n_oph
      integer(c_int32_t), intent(out) :: n_h
      integer(c_int32_t), dimension(*), intent(out) :: h_desc
      integer(c_int32_t), value :: cap
    end function nsx_ens_layout

    integer(c_int) function nsx_ens_values(h, values) bind(C, name='nsx_ens_values')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), dimension(*), intent(out) :: values
    end function nsx_ens_values

    ! templates, bases: arrays of c_ptr (bases(t) = c_loc of a P x K_t row-major array);
    ! theta: S x P row-major (theta(i, s) in Fortran order); cfg may be c_null_ptr
    integer(c_int) function nsx_ens_create(templates, n_t, s, p, theta, bases, precision, cfg, b) &
        bind(C, name='nsx_ens_create')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), dimension(*), intent(in) :: templates, bases
      integer(c_int32_t), value :: n_t, s, p
      real(c_double), dimension(*), intent(in) :: theta
      integer(c_int), value :: precision
      type(c_ptr), value :: cfg
      type(c_ptr), intent(out) :: b
    end function nsx_ens_create

    integer(c_int) function nsx_ens_update(b, theta) bind(C, name='nsx_ens_update')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      real(c_double), dimension(*), intent(in) :: theta
    end function nsx_ens_update

    type(c_ptr) function nsx_last_error() bind(C, name='nsx_last_error')
      import :: c_ptr
    end function nsx_last_error

    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function c_strlen
  end interface

contains

  ! blank-trimmed, NUL-terminated copy of a Fortran string
  function nsx_cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(adjustl(s)) + 1) :: c
    c = trim(adjustl(s)) // c_null_char
  end function nsx_cstr

  ! the library's last error message as a Fortran string
  function nsx_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = nsx_last_error()
    n = int(c_strlen(p))
    allocate(character(len=n) :: msg)
    if (n == 0) return
    call c_f_pointer(p, chars, [n])
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function nsx_error_message

end module nsx_hv_api
