! Test infrastructure only (never linked into the product).
!
! Runs the reference's own jacobi_diis (jacobi_diis.f90, compiled unmodified
! from its source) on a dense linear system (D + A) x = rhs with callbacks
! written here: matvec y = A x, dm1vec y = x / d and the user norm
! u_norm = sqrt(sum_i x_i^2 / i / n).  This pins the solver's plugin
! interface (jacobi_diis.f90:43-55): every norm choice, lprint, DIIS or plain
! Jacobi, and the iteration count on failure.
!
! usage: jd_dump <input.bin> <output.bin>
!   input (stream, little-endian): int32 n, diis_max, norm, n_iter, lprint;
!   real*8 tol; then A (n,n) column-major, d(n), rhs(n), x0(n)
!   output: int32 n_iter, ok; real*8 x(n).  The residual lines go to stdout.
module jd_data
  implicit none
  real*8, allocatable :: a(:,:), d(:)
end module jd_data

subroutine dense_mv(n, x, y)
  use jd_data
  implicit none
  integer, intent(in) :: n
  real*8, intent(in) :: x(n)
  real*8, intent(inout) :: y(n)
  integer :: i, j
  do i = 1, n
    y(i) = 0.d0
    do j = 1, n
      y(i) = y(i) + a(i,j)*x(j)
    end do
  end do
end subroutine dense_mv

subroutine dense_dm1(n, x, y)
  use jd_data
  implicit none
  integer, intent(in) :: n
  real*8, intent(in) :: x(n)
  real*8, intent(inout) :: y(n)
  y = x/d
end subroutine dense_dm1

real*8 function dense_norm(n, x)
  implicit none
  integer, intent(in) :: n
  real*8, intent(in) :: x(n)
  integer :: i
  dense_norm = 0.d0
  do i = 1, n
    dense_norm = dense_norm + x(i)*x(i)/dble(i)
  end do
  dense_norm = sqrt(dense_norm/dble(n))
end function dense_norm

program jd_dump
  use jd_data
  implicit none
  integer :: n, diis_max, norm, n_iter, lprint
  real*8 :: tol
  real*8, allocatable :: rhs(:), x(:)
  logical :: ok
  integer :: iok
  character(len=512) :: fin, fout
  external :: dense_mv, dense_dm1
  real*8, external :: dense_norm
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(unit=10, file=trim(fin), form='unformatted', access='stream', status='old')
  read(10) n, diis_max, norm, n_iter, lprint
  read(10) tol
  allocate(a(n,n), d(n), rhs(n), x(n))
  read(10) a, d, rhs, x
  close(10)
  ok = .false.
  call jacobi_diis(n, lprint, diis_max, norm, tol, rhs, x, n_iter, ok, dense_mv, dense_dm1, &
                   dense_norm)
  iok = 0
  if (ok) iok = 1
  open(unit=11, file=trim(fout), form='unformatted', access='stream', status='replace')
  write(11) n_iter, iok
  write(11) x
  close(11)
end program jd_dump
