! Drop-in driver: reads Input.txt (or the file named on the command line)
! and runs the reference driver's calling sequence (main.f90:124-282) --
! ddinit, mkrhs, the direct solve and energy, then with igrad the adjoint
! solve, forces_dd, zeta and both field contractions -- on the GPU through
! the ddcosmo module of ddcosmo_gpu.f90.  Output format follows the
! reference's log: energy in kcal/mol, forces in hartree/bohr.
program ddx_main
  use ddcosmo
  implicit none
  integer :: i, ii, isph, ig, n, nargs, ios
  real*8 :: tobohr, esolv, xx(1)
  real*8, parameter :: toang = 0.52917721092d0, tokcal = 627.509469d0
  real*8, allocatable :: x(:), y(:), z(:), rvdw(:), charge(:)
  real*8, allocatable :: phi(:), psi(:,:), sigma(:,:), s(:,:), fx(:,:), zeta(:), ef(:,:)
  real*8, allocatable :: zpad(:)
  character(len=512) :: fname

  fname = 'Input.txt'
  nargs = command_argument_count()
  if (nargs >= 1) call get_command_argument(1, fname)
  open(unit=100, file=trim(fname), form='formatted', access='sequential', status='old', &
       iostat=ios)
  if (ios /= 0) then
    write(*,*) 'ddx_main: cannot open ', trim(fname)
    stop
  end if
  read(100,*) iprint
  read(100,*) nproc
  read(100,*) lmax
  read(100,*) ngrid
  read(100,*) iconv
  read(100,*) igrad
  read(100,*) eps
  read(100,*) eta
  read(100,*) n
  allocate(x(n), y(n), z(n), rvdw(n), charge(n))
  do i = 1, n
    read(100,*) charge(i), x(i), y(i), z(i), rvdw(i)
  end do
  close(100)
  ! angstrom to bohr (main.f90:149-153)
  tobohr = 1.0d0/toang
  x = x*tobohr ; y = y*tobohr ; z = z*tobohr ; rvdw = rvdw*tobohr

  call ddinit(n, x, y, z, rvdw)
  allocate(phi(ncav), psi(nylm,n), sigma(nylm,n))
  call mkrhs(n, charge, x, y, z, ncav, ccav, phi, nylm, psi)
  call cosmo(.false., .true., phi, xx, psi, sigma, esolv)
  if (iprint >= 3) call prtsph('solution to the ddCOSMO equation', nsph, 0, sigma)
  write(6, '(1x,a,f14.6)') 'ddcosmo electrostatic solvation energy (kcal/mol):', esolv*tokcal

  if (igrad == 1) then
    write(6,*)
    allocate(s(nylm,n), fx(3,n), zeta(ncav), ef(3,max(n,ncav)))
    call cosmo(.true., .false., xx, xx, psi, s, esolv)
    if (iprint >= 3) call prtsph('solution to the ddCOSMO adjoint equation', nsph, 0, s)
    call forces_dd(n, phi, sigma, s, fx)
    call ddmkzeta(s, zeta)
    ! (main.f90:214 prints zeta(ncav) as an (ngrid, nsph) table, reading
    ! past its end; here the positions beyond ncav print as zero)
    if (iprint >= 4) then
      allocate(zpad(ngrid*nsph))
      zpad = zero
      zpad(1:ncav) = zeta
      call ptcart('zeta', nsph, 0, zpad)
      deallocate(zpad)
    end if
    ! the solute's field at the cavity points contracted with zeta
    call efld(n, charge, csph, ncav, ccav, ef)
    ii = 0
    do isph = 1, nsph
      do ig = 1, ngrid
        if (ui(ig,isph) > zero) then
          ii = ii + 1
          fx(:,isph) = fx(:,isph) - zeta(ii)*ef(:,ii)
        end if
      end do
    end do
    ! the field of zeta at the nuclei times the charges
    call efld(ncav, zeta, ccav, n, csph, ef)
    do isph = 1, nsph
      fx(:,isph) = fx(:,isph) - ef(:,isph)*charge(isph)
    end do
    if (iprint >= 1) then
      write(iout, '(1x,"ddCOSMO forces (atomic units):",/,1x," atom",15x,"x",15x,"y",15x,"z")')
      do isph = 1, nsph
        write(6, '(1x,i5,3f16.8)') isph, fx(:,isph)
      end do
    end if
    deallocate(s, fx, zeta, ef)
  end if
  deallocate(x, y, z, rvdw, charge, phi, psi, sigma)
  call memfree
end program ddx_main
