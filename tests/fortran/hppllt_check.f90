! Caller-side check of the Fortran hppllt module, written the way a program
! using the reference module calls it (example.f90:63-145 drives the same
! sequence).  Matrix: 2D 5-point Laplacian on a k x k grid (4 on the
! diagonal, -1 off it), lower CCS with the diagonal first in each column.
!
! usage: hppllt_check k nb max_zero mode
!   mode "analyze": analysis and getters only (no GPU needed)
!   mode "solve":   also factorize + solve (1 and 3 right-hand sides, FP64
!                   and FP32 + refinement) and check x against x_true
! Prints "key value" lines; stops with code 1 on a failed check.
program hppllt_check
  use hppllt
  implicit none
  integer :: k, nb, max_zero, n, nnz, i, j, col, p, r
  integer, allocatable :: ccs_col(:), ccs_row(:), perm(:)
  double precision, pointer, contiguous :: a(:)
  double precision, allocatable :: xt(:), b(:), bm(:, :), xm(:, :)
  character(len=32) :: arg, mode
  double precision :: err
  double precision :: eta3(3), res3(3), ferr3(3)

  call get_command_argument(1, arg); read (arg, *) k
  call get_command_argument(2, arg); read (arg, *) nb
  call get_command_argument(3, arg); read (arg, *) max_zero
  call get_command_argument(4, mode)
  n = k * k

  ! lower CCS: column (i, j) -> rows (i, j), (i+1, j), (i, j+1)
  allocate(ccs_col(n + 1), ccs_row(3 * n), a(3 * n))
  nnz = 0
  ccs_col(1) = 1
  do j = 1, k
    do i = 1, k
      col = i + (j - 1) * k
      nnz = nnz + 1; ccs_row(nnz) = col; a(nnz) = 4d0
      if (i < k) then
        nnz = nnz + 1; ccs_row(nnz) = col + 1; a(nnz) = -1d0
      endif
      if (j < k) then
        nnz = nnz + 1; ccs_row(nnz) = col + k; a(nnz) = -1d0
      endif
      ccs_col(col + 1) = nnz + 1
    enddo
  enddo

  call hppllt_init([nb, max_zero, 0, 0])
  call hppllt_analyze(ccs_col, ccs_row(1:nnz), n)
  if (hppllt_get_status() /= 0) stop 1
  print '(a,i0)', "n ", n
  print '(a,i0)', "nnz ", nnz
  print '(a,i0)', "cost ", hppllt_get_cost()
  print '(a,i0)', "nonzero ", hppllt_get_nonzero()
  print '(a,i0)', "array_size_1 ", hppllt_get_array_size(1)
  print '(a,i0)', "array_size_2 ", hppllt_get_array_size(2)
  print '(a,i0)', "array_size_3 ", hppllt_get_array_size(3)

  if (trim(mode) == "solve") then
    allocate(xt(n), b(n))
    do i = 1, n
      xt(i) = dble(mod(7 * i, 13)) - 6d0
    enddo
    call matvec(xt, b)
    call hppllt_factorize(a)
    if (hppllt_get_status() /= 0) stop 1
    call hppllt_solve(b)
    if (hppllt_get_status() /= 0) stop 1
    err = maxval(abs(b - xt)) / maxval(abs(xt))
    print '(a,es10.3)', "fp64_x_err ", err
    if (err > 1d-10) stop 1

    allocate(bm(n, 3), xm(n, 3))
    do r = 1, 3
      do i = 1, n
        xm(i, r) = dble(mod(5 * i + r, 11)) - 5d0
      enddo
      call matvec(xm(:, r), bm(:, r))
    enddo
    call hppllt_solve_multi(bm)
    if (hppllt_get_status() /= 0) stop 1
    err = maxval(abs(bm - xm)) / maxval(abs(xm))
    print '(a,es10.3)', "multi_x_err ", err
    if (err > 1d-10) stop 1
    call hppllt_get_solve_errors(eta3, res3, ferr3)
    if (hppllt_get_status() /= 0) stop 1
    print '(a,3es10.3)', "multi_backward_error ", eta3
    print '(a,3es10.3)', "multi_forward_error ", ferr3
    if (maxval(eta3) > 1d-14 .or. maxval(ferr3) > 1d-10 .or. hppllt_get_backward_error() /= maxval(eta3)) stop 1

    call hppllt_set_precision(1)
    call hppllt_factorize(a)
    if (hppllt_get_status() /= 0) stop 1
    call matvec(xt, b)
    call hppllt_solve(b)
    if (hppllt_get_status() /= 0) stop 1
    err = maxval(abs(b - xt)) / maxval(abs(xt))
    print '(a,es10.3)', "fp32ir_x_err ", err
    print '(a,es10.3)', "fp32ir_backward_error ", hppllt_get_backward_error()
    if (err > 1d-10 .or. hppllt_get_backward_error() > 1d-14) stop 1
    print '(a,f10.6)', "factorize_time ", hppllt_get_time_info(3)
    print '(a,f10.6)', "solve_time ", hppllt_get_time_info(4)
  endif

  ! use_metis = 1 with a caller ordering: red-black (grid points with i + j
  ! even first, then the odd ones, each in natural order)
  call hppllt_finalize()
  call hppllt_init([nb, max_zero, 0, 1])
  allocate(perm(n))
  p = 0
  do r = 0, 1
    do col = 1, n
      i = mod(col - 1, k) + 1
      j = (col - 1) / k + 1
      if (mod(i + j, 2) == r) then
        p = p + 1
        perm(p) = col
      endif
    enddo
  enddo
  call hppllt_set_ordering(perm)
  call hppllt_analyze(ccs_col, ccs_row(1:nnz), n)
  if (hppllt_get_status() /= 0) stop 1
  print '(a,i0)', "perm_cost ", hppllt_get_cost()
  if (trim(mode) == "solve") then
    call matvec(xt, b)
    call hppllt_factorize(a)
    call hppllt_solve(b)
    if (hppllt_get_status() /= 0) stop 1
    err = maxval(abs(b - xt)) / maxval(abs(xt))
    print '(a,es10.3)', "perm_x_err ", err
    if (err > 1d-10) stop 1
  endif
  call hppllt_finalize()

  ! use_metis = 1 without a caller ordering: the library's own ordering
  call hppllt_init([nb, max_zero, 0, 1])
  call hppllt_analyze(ccs_col, ccs_row(1:nnz), n)
  if (hppllt_get_status() /= 0) stop 1
  print '(a,i0)', "nd_cost ", hppllt_get_cost()
  if (trim(mode) == "solve") then
    call matvec(xt, b)
    call hppllt_factorize(a)
    call hppllt_solve(b)
    if (hppllt_get_status() /= 0) stop 1
    err = maxval(abs(b - xt)) / maxval(abs(xt))
    print '(a,es10.3)', "nd_x_err ", err
    if (err > 1d-10) stop 1
  endif
  call hppllt_finalize()
  print '(a)', "ok"

contains

  ! y = A x with A symmetric, from the lower CCS.
  subroutine matvec(x, y)
    double precision, intent(in) :: x(:)
    double precision, intent(out) :: y(:)
    integer :: c, e, row
    y = 0d0
    do c = 1, n
      do e = ccs_col(c), ccs_col(c + 1) - 1
        row = ccs_row(e)
        y(row) = y(row) + a(e) * x(c)
        if (row /= c) y(c) = y(c) + a(e) * x(row)
      enddo
    enddo
  end subroutine

end program
