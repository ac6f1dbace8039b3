! Fortran module hppllt over the MI355X numeric phase (libmfdag, C-ABI in
! include/mfdag.h), bound through ISO_C_BINDING.
!
! The public names, argument declarations and option layout are those of the
! reference module (to-mo-ki/multifrontal-DAG, src/main/hppllt.f90:8-9,
! options src/main/options_m.f90:3-7), so a caller of the reference links
! against this module unchanged.  The analysis runs on the host, the
! factorization and the solves on one gfx950 GPU.
!
! Differences, all additive:
!   - with options(4) = use_metis = 1 the fill-reducing ordering comes from
!     the library's graph nested dissection (mfd_order_nd) where the
!     reference calls METIS_NodeND, unless hppllt_set_ordering(perm) (new ->
!     old, 1-based) supplies one;
!   - hppllt_set_precision(p): 0 = FP64 (default), 1 = FP32 fronts + FP64
!     iterative refinement, 2 / 3 = FP16 / BF16 updates of the diagonally
!     scaled fronts + FP64 iterative refinement;
!   - hppllt_get_solve_errors(eta, res2, ferr): per right-hand side backward
!     error, ||b - Ax||_2/||x||_2 and forward error estimate of the last solve;
!   - hppllt_solve_multi(b(:,:)) solves several right-hand sides at once;
!   - hppllt_get_status() returns the last library return code (0 = ok, > 0 =
!     1-based global column of a non-positive pivot, < 0 = error); the
!     reference has no error reporting, so errors are also printed;
!   - hppllt_get_nonzero8() returns the int64 count (hppllt_get_nonzero keeps
!     the reference's default-integer result and saturates);
!   - hppllt_finalize releases the device state, after which hppllt_init may
!     be called again; ccs_row is copied by analyze, not retained.
module hppllt
  use iso_c_binding
  implicit none
  private

  public :: hppllt_init, hppllt_analyze, hppllt_factorize, hppllt_solve, hppllt_finalize
  public :: hppllt_get_time_info, hppllt_get_cost, hppllt_get_nonzero, hppllt_get_array_size
  public :: hppllt_solve_multi, hppllt_set_ordering, hppllt_set_precision, hppllt_get_status
  public :: hppllt_get_nonzero8, hppllt_get_backward_error, hppllt_get_solve_errors

  integer, parameter :: NUM_OPTIONS = 4
  integer, parameter :: OPTION_NB = 1, OPTION_MAX_ZERO = 2, OPTION_USE_STARPU = 3, OPTION_USE_METIS = 4
  integer(c_int), parameter :: FLAG_GRAPH = 1, FLAG_ORDER_ND = 4
  integer(c_int), parameter :: ERR_NODEVICE = -7

  type(c_ptr), save :: handle = c_null_ptr
  integer, allocatable, save :: options(:)
  integer(c_int), allocatable, target, save :: ordering(:)
  integer(c_int), save :: precision = 0
  integer(c_int), save :: status = 0
  integer(c_int), save :: last_nrhs = 0

  type, bind(C) :: mfd_stats
    real(c_double) :: analyze_ms, assemble_ms, factor_ms, solve_ms
    real(c_double) :: factor_gflops
    integer(c_int) :: precision, refine_iters
    real(c_double) :: backward_error, refine_residual
    integer(c_int64_t) :: front_bytes, device_bytes
    integer(c_int) :: num_levels, num_launches_factor, num_launches_solve
  end type

  interface
    integer(c_int) function mfd_create(h, nb, flags) bind(C, name="mfd_create")
      import :: c_ptr, c_int
      type(c_ptr) :: h
      integer(c_int), value :: nb, flags
    end function
    subroutine mfd_destroy(h) bind(C, name="mfd_destroy")
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
    integer(c_int) function mfd_analyze(h, n, ccs_col, ccs_row, max_zero, perm) bind(C, name="mfd_analyze")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: n, max_zero
      integer(c_int), intent(in) :: ccs_col(*), ccs_row(*)
      type(c_ptr), value :: perm
    end function
    integer(c_int) function mfd_factorize(h, a, prec) bind(C, name="mfd_factorize")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: a(*)
      integer(c_int), value :: prec
    end function
    integer(c_int) function mfd_solve(h, b, nrhs, ldb) bind(C, name="mfd_solve")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double) :: b(*)
      integer(c_int), value :: nrhs, ldb
    end function
    integer(c_int64_t) function mfd_get_cost(h) bind(C, name="mfd_get_cost")
      import :: c_ptr, c_int64_t
      type(c_ptr), value :: h
    end function
    integer(c_int64_t) function mfd_get_nonzero(h) bind(C, name="mfd_get_nonzero")
      import :: c_ptr, c_int64_t
      type(c_ptr), value :: h
    end function
    integer(c_int64_t) function mfd_get_array_size(h, which) bind(C, name="mfd_get_array_size")
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: which
    end function
    real(c_double) function mfd_get_time_info(h, id) bind(C, name="mfd_get_time_info")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: id
    end function
    integer(c_int) function mfd_get_stats(h, st) bind(C, name="mfd_get_stats")
      import :: c_ptr, c_int, mfd_stats
      type(c_ptr), value :: h
      type(mfd_stats) :: st
    end function
    integer(c_int) function mfd_get_solve_errors(h, nrhs, eta, res2, ferr) bind(C, name="mfd_get_solve_errors")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nrhs
      type(c_ptr), value :: eta, res2, ferr
    end function
    type(c_ptr) function mfd_error_string(code) bind(C, name="mfd_error_string")
      import :: c_ptr, c_int
      integer(c_int), value :: code
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Records rc; prints the library's message for an error or a failed pivot.
  subroutine report(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: length
    status = rc
    if (rc == 0) return
    if (rc > 0) then
      print '(a,a,i0)', what, ": matrix is not positive definite at column ", rc
      return
    endif
    length = int(c_strlen(mfd_error_string(rc)))
    call c_f_pointer(mfd_error_string(rc), msg, [length])
    print '(a,a,i0,a,*(a))', what, ": error ", rc, " ", msg
  end subroutine

  ! hppllt.f90:13-22: options = [nb, max_zero, use_starpu, use_metis].
  subroutine hppllt_init(input_options)
    integer :: input_options(:)
    integer(c_int) :: rc
    if (size(input_options) /= NUM_OPTIONS) then
      print *, "incorrect size of options"
      return
    endif
    if (allocated(options)) then
      print *, "hppllt_init: already initialised (call hppllt_finalize first)"
      return
    endif
    allocate(options, source=input_options)
    if (options(OPTION_USE_METIS) == 1) then
      rc = mfd_create(handle, int(options(OPTION_NB), c_int), ior(FLAG_GRAPH, FLAG_ORDER_ND))
    else
      rc = mfd_create(handle, int(options(OPTION_NB), c_int), FLAG_GRAPH)
    endif
    ! Without a GPU the handle still serves the analysis and the getters.
    if (rc == ERR_NODEVICE) then
      status = rc
    else
      call report(rc, "hppllt_init")
    endif
  end subroutine

  ! Ordering for use_metis = 1 in place of the library's own: perm(k) is the
  ! original index of the k-th unknown in the new order (1-based), the
  ! orientation of reordering_perm in the reference (hppllt.f90:47-51).
  subroutine hppllt_set_ordering(perm)
    integer, intent(in) :: perm(:)
    if (allocated(ordering)) deallocate(ordering)
    allocate(ordering(size(perm)))
    ordering = int(perm, c_int)
  end subroutine

  subroutine hppllt_set_precision(p)
    integer, intent(in) :: p
    precision = int(p, c_int)
  end subroutine

  ! hppllt.f90:24-81: lower CCS (1-based, diagonal first in each column).
  subroutine hppllt_analyze(ccs_col, ccs_row, n)
    integer, contiguous :: ccs_col(:), ccs_row(:)
    integer, intent(in) :: n
    type(c_ptr) :: perm
    integer(c_int) :: rc
    if (.not. allocated(options)) then
      print *, "hppllt_analyze: hppllt_init has not been called"
      return
    endif
    perm = c_null_ptr
    if (options(OPTION_USE_METIS) == 1 .and. allocated(ordering)) then
      if (size(ordering) /= n) then
        print *, "hppllt_analyze: ordering has the wrong length"
        status = -1
        return
      endif
      perm = c_loc(ordering)
    endif
    rc = mfd_analyze(handle, int(n, c_int), ccs_col, ccs_row, int(options(OPTION_MAX_ZERO), c_int), perm)
    call report(rc, "hppllt_analyze")
  end subroutine

  ! hppllt.f90:123-150: values in the caller's CCS order.
  subroutine hppllt_factorize(a)
    double precision, pointer, contiguous :: a(:)
    call report(mfd_factorize(handle, a, precision), "hppllt_factorize")
  end subroutine

  ! hppllt.f90:152-173: b is overwritten with x.
  subroutine hppllt_solve(b)
    double precision, contiguous :: b(:)
    call report(mfd_solve(handle, b, 1_c_int, int(size(b), c_int)), "hppllt_solve")
    last_nrhs = 1
  end subroutine

  ! Several right-hand sides: b(n, nrhs) is overwritten with X.
  subroutine hppllt_solve_multi(b)
    double precision, contiguous :: b(:, :)
    call report(mfd_solve(handle, b, int(size(b, 2), c_int), int(size(b, 1), c_int)), "hppllt_solve_multi")
    last_nrhs = int(size(b, 2), c_int)
  end subroutine

  ! hppllt.f90:175-184; frees the device state.
  subroutine hppllt_finalize()
    if (c_associated(handle)) call mfd_destroy(handle)
    handle = c_null_ptr
    if (allocated(options)) deallocate(options)
    if (allocated(ordering)) deallocate(ordering)
    precision = 0
  end subroutine

  ! hppllt.f90:186-189, timer ids of time_list_m.f90:3-10, seconds.
  double precision function hppllt_get_time_info(num) result(time)
    integer :: num
    time = mfd_get_time_info(handle, int(num, c_int))
  end function

  ! hppllt.f90:191-194 (cost_calculator_m.f90:9-20).
  integer(8) function hppllt_get_cost() result(cost)
    cost = mfd_get_cost(handle)
  end function

  ! hppllt.f90:196-199 (cost_calculator_m.f90:22-33).
  integer function hppllt_get_nonzero() result(nonzero)
    integer(c_int64_t) :: v
    v = mfd_get_nonzero(handle)
    nonzero = int(min(v, int(huge(nonzero), c_int64_t)))
  end function

  integer(8) function hppllt_get_nonzero8() result(nonzero)
    nonzero = mfd_get_nonzero(handle)
  end function

  ! hppllt.f90:201-211: 1 supernode, 2 work, 3 border array size.
  integer(8) function hppllt_get_array_size(num) result(array_size)
    integer :: num
    array_size = mfd_get_array_size(handle, int(num, c_int))
  end function

  integer function hppllt_get_status() result(rc)
    rc = status
  end function

  ! Normwise backward error of the last solve (max over right-hand sides),
  ! in either precision form; -1 before a solve.
  double precision function hppllt_get_backward_error() result(eta)
    double precision, allocatable, target :: e(:)
    integer(c_int) :: rc
    eta = -1d0
    if (last_nrhs < 1) return
    allocate(e(last_nrhs))
    rc = mfd_get_solve_errors(handle, last_nrhs, c_loc(e), c_null_ptr, c_null_ptr)
    if (rc == 0) eta = maxval(e)
  end function

  ! Errors of the last solve per right-hand side (each array sized to that
  ! solve's number of right-hand sides): eta the normwise backward error,
  ! res2 the reference driver's ||b - Ax||_2 / ||x||_2 (example.f90:147-180),
  ! ferr a forward error estimate (one more solve with the current factor).
  subroutine hppllt_get_solve_errors(eta, res2, ferr)
    double precision, contiguous, intent(out), target :: eta(:), res2(:), ferr(:)
    if (size(eta) /= last_nrhs .or. size(res2) /= last_nrhs .or. size(ferr) /= last_nrhs) then
      print *, "hppllt_get_solve_errors: arrays must hold one entry per right-hand side of the last solve"
      status = -1
      return
    endif
    call report(mfd_get_solve_errors(handle, last_nrhs, c_loc(eta), c_loc(res2), c_loc(ferr)), &
                "hppllt_get_solve_errors")
  end subroutine

end module
