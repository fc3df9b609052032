! Test program for the drop-in module (module ddcosmo over libddx): the
! reference's solver interface jacobi_diis (jacobi_diis.f90:43-55) and mkrhs
! with coordinates of the caller's choosing (mkrhs.f90:1).
!
! usage: jd_module dense <input.bin> <output.bin>
!          the dense system of oracle/jd_dump.f90 (same files), solved by
!          the module's jacobi_diis with the same callbacks
!        jd_module module <Input.txt deck> <output.bin>
!          ddinit + mkrhs + cosmo on the deck; then the adjoint system
!          L* S = psi by jacobi_diis with the module's lstarx, ldm1x, hnorm
!          (the device path) and with a caller's wrapper of lstarx (the
!          host-callback path); and mkrhs at coordinates moved by 0.1 bohr.
!          output: int32 it_fast, ok_fast, it_wrap, ok_wrap, it_cosmo;
!          real*8 S_fast, S_wrap, S_cosmo (nylm, n), phi_moved (ncav)
module jdm_data
  implicit none
  real*8, allocatable :: a(:,:), d(:)
end module jdm_data

subroutine dense_mv(n, x, y)
  use jdm_data
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
  use jdm_data
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

! a caller's own operator: the module's lstarx behind an external name
subroutine my_lstarx(n, x, y)
  use ddcosmo, only: lstarx
  implicit none
  integer, intent(in) :: n
  real*8, intent(in) :: x(n)
  real*8, intent(inout) :: y(n)
  call lstarx(n, x, y)
end subroutine my_lstarx

program jd_module
  use ddcosmo
  use jdm_data
  implicit none
  character(len=512) :: mode, fin, fout
  integer :: n, dmax, nrm, n_iter, lp, i, it1, it2, isph
  integer :: ok1i, ok2i
  real*8 :: tol, esolv, xx(1)
  logical :: ok, ok2
  real*8, allocatable :: rhs(:), x(:)
  real*8, allocatable :: xs(:), ys(:), zs(:), rvdw(:), charge(:), phi(:), psi(:,:), sigma(:,:)
  real*8, allocatable :: s1(:,:), s2(:,:), s0(:,:), phim(:), psim(:,:)
  real*8, parameter :: toang = 0.52917721092d0
  external :: dense_mv, dense_dm1, my_lstarx
  real*8, external :: dense_norm
  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  if (trim(mode) == 'dense') then
    open(unit=10, file=trim(fin), form='unformatted', access='stream', status='old')
    read(10) n, dmax, nrm, n_iter, lp
    read(10) tol
    allocate(a(n,n), d(n), rhs(n), x(n))
    read(10) a, d, rhs, x
    close(10)
    ok = .false.
    call jacobi_diis(n, lp, dmax, nrm, tol, rhs, x, n_iter, ok, dense_mv, dense_dm1, dense_norm)
    ok1i = 0
    if (ok) ok1i = 1
    open(unit=11, file=trim(fout), form='unformatted', access='stream', status='replace')
    write(11) n_iter, ok1i
    write(11) x
    close(11)
  else
    open(unit=10, file=trim(fin), form='formatted', status='old')
    read(10,*) iprint
    read(10,*) nproc
    read(10,*) lmax
    read(10,*) ngrid
    read(10,*) iconv
    read(10,*) igrad
    read(10,*) eps
    read(10,*) eta
    read(10,*) n
    allocate(xs(n), ys(n), zs(n), rvdw(n), charge(n))
    do i = 1, n
      read(10,*) charge(i), xs(i), ys(i), zs(i), rvdw(i)
    end do
    close(10)
    xs = xs/toang ; ys = ys/toang ; zs = zs/toang ; rvdw = rvdw/toang
    iprint = 0
    call ddinit(n, xs, ys, zs, rvdw)
    allocate(phi(ncav), psi(nylm,n), sigma(nylm,n), s0(nylm,n), s1(nylm,n), s2(nylm,n))
    call mkrhs(n, charge, xs, ys, zs, ncav, ccav, phi, nylm, psi)
    call cosmo(.false., .true., phi, xx, psi, sigma, esolv)
    call cosmo(.true., .false., xx, xx, psi, s0, esolv)
    tol = 10.0d0**(-iconv)
    ! the device path
    call ldm1x(nylm*nsph, psi, s1)
    it1 = 200
    call jacobi_diis(nylm*nsph, 0, ndiis, 4, tol, psi, s1, it1, ok, lstarx, ldm1x, hnorm)
    ok1i = 0
    if (ok) ok1i = 1
    ! the host-callback path with the caller's operator
    call ldm1x(nylm*nsph, psi, s2)
    it2 = 200
    call jacobi_diis(nylm*nsph, 0, ndiis, 4, tol, psi, s2, it2, ok2, my_lstarx, ldm1x, hnorm)
    ok2i = 0
    if (ok2) ok2i = 1
    ! mkrhs at moved coordinates (the cavity points stay)
    allocate(phim(ncav), psim(nylm,n))
    call mkrhs(n, charge, xs + 0.1d0, ys - 0.05d0, zs, ncav, ccav, phim, nylm, psim)
    open(unit=11, file=trim(fout), form='unformatted', access='stream', status='replace')
    write(11) it1, ok1i, it2, ok2i, 0
    write(11) s1, s2, s0, phim
    close(11)
    isph = 0
    call memfree
  end if
end program jd_module
