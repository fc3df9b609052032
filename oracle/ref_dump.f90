! Test infrastructure only (never linked into the product).
!
! Drives the reference ddCOSMO module, compiled from its own sources, through
! the same calling sequence as the reference driver (main.f90:124-282) and
! writes every intermediate in full binary FP64 so that fixtures carry more
! than the 6-8 printed decimals of the shipped logs.
!
! usage: ref_dump <Input.txt-format file> <output prefix> [probe]
!   writes <prefix>.hdr (key value lines) and <prefix>.bin (raw little-endian
!   arrays, Fortran column-major, in the order listed in the header).
!   With "probe" it also applies lx / lstarx / hnorm to a fixed test vector.
!
program ref_dump
use ddcosmo
implicit none
integer :: i, ii, isph, ig, n, nargs, ios, k
real*8 :: tobohr, esolv, xx(1)
real*8, parameter :: toang=0.52917721092d0
real*8, allocatable :: x(:), y(:), z(:), rvdw(:), charge(:)
real*8, allocatable :: phi(:), psi(:,:), sigma(:,:), s(:,:), g(:,:), rhs(:,:)
real*8, allocatable :: fx(:,:), fxdd(:,:), zeta(:), ef(:,:), xi(:,:)
real*8, allocatable :: v(:,:), lv(:,:), lsv(:,:)
real*8 :: hv, edum
character(len=512) :: fin, prefix, mode
logical :: probe
external :: lx, lstarx
real*8, external :: hnorm
!
nargs = command_argument_count()
call get_command_argument(1, fin)
call get_command_argument(2, prefix)
probe = .false.
if (nargs.ge.3) then
  call get_command_argument(3, mode)
  probe = (trim(mode).eq.'probe')
end if
!
open (unit=100,file=trim(fin),form='formatted',access='sequential',status='old',iostat=ios)
if (ios.ne.0) stop 'ref_dump: cannot open input'
read(100,*) iprint
read(100,*) nproc
read(100,*) lmax
read(100,*) ngrid
read(100,*) iconv
read(100,*) igrad
read(100,*) eps
read(100,*) eta
read(100,*) n
allocate (x(n),y(n),z(n),rvdw(n),charge(n))
do i = 1, n
  read(100,*) charge(i), x(i), y(i), z(i), rvdw(i)
end do
close(100)
tobohr = 1.0d0/toang
x = x*tobohr ; y = y*tobohr ; z = z*tobohr ; rvdw = rvdw*tobohr
! keep the header and the per-iteration residual lines, nothing bulkier
iprint = min(iprint,1)
if (iprint.lt.1) iprint = 1
!
call ddinit(n,x,y,z,rvdw)
allocate (phi(ncav),psi(nylm,n),sigma(nylm,n),g(ngrid,n),rhs(nylm,n))
call mkrhs(n,charge,x,y,z,ncav,ccav,phi,nylm,psi)
! the right-hand side exactly as cosmo assembles it (cosmo.f90:111-116)
call wghpot(phi,g)
do isph = 1, nsph
  call intrhs(isph,g(:,isph),rhs(:,isph))
end do
call cosmo(.false., .true., phi, xx, psi, sigma, esolv)
!
allocate (s(nylm,n),fx(3,n),fxdd(3,n),zeta(ncav),ef(3,max(n,ncav)),xi(ngrid,n))
s = 0.d0 ; fx = 0.d0 ; fxdd = 0.d0 ; zeta = 0.d0 ; xi = 0.d0
if (igrad.eq.1) then
  call cosmo(.true., .false., xx, xx, psi, s, edum)
  do isph = 1, nsph
    do ig = 1, ngrid
      xi(ig,isph) = dot_product(s(:,isph),basis(:,ig))
    end do
  end do
  call forces_dd(n,phi,sigma,s,fx)
  fxdd = fx
  call ddmkzeta(s,zeta)
  call efld(n,charge,csph,ncav,ccav,ef)
  ii = 0
  do isph = 1, nsph
    do ig = 1, ngrid
      if (ui(ig,isph).gt.zero) then
        ii = ii + 1
        fx(:,isph) = fx(:,isph) - zeta(ii)*ef(:,ii)
      end if
    end do
  end do
  call efld(ncav,zeta,ccav,n,csph,ef)
  do isph = 1, nsph
    fx(:,isph) = fx(:,isph) - ef(:,isph)*charge(isph)
  end do
end if
!
if (probe) then
  allocate (v(nylm,n),lv(nylm,n),lsv(nylm,n))
  k = 0
  do isph = 1, nsph
    do i = 1, nylm
      k = k + 1
      v(i,isph) = sin(0.37d0*dble(k)) + 0.25d0*cos(1.3d0*dble(isph))
    end do
  end do
  call lx(nylm*nsph, v, lv)
  call lstarx(nylm*nsph, v, lsv)
  hv = hnorm(nylm*nsph, v)
end if
!
open (unit=200,file=trim(prefix)//'.hdr',form='formatted',status='replace')
write(200,'(a,1x,i12)') 'nsph', nsph
write(200,'(a,1x,i12)') 'ngrid', ngrid
write(200,'(a,1x,i12)') 'lmax', lmax
write(200,'(a,1x,i12)') 'nylm', nylm
write(200,'(a,1x,i12)') 'ncav', ncav
write(200,'(a,1x,i12)') 'igrad', igrad
write(200,'(a,1x,i12)') 'iconv', iconv
write(200,'(a,1x,i12)') 'nnl', inl(nsph+1)-1
write(200,'(a,1x,es25.17)') 'eps', eps
write(200,'(a,1x,es25.17)') 'eta', eta
write(200,'(a,1x,es25.17)') 'esolv', esolv
if (probe) write(200,'(a,1x,es25.17)') 'hnorm_probe', hv
write(200,'(a)') 'arrays grid w basis facs facl inl nl fi ui zi ccav phi psi rhs sigma s xi fxdd zeta fx probe'
close(200)
!
open (unit=201,file=trim(prefix)//'.bin',form='unformatted',access='stream',status='replace')
write(201) grid, w, basis, facs, facl
write(201) inl, nl(1:inl(nsph+1)-1)
write(201) fi, ui
if (igrad.eq.1) then
  write(201) zi
end if
write(201) ccav, phi, psi, rhs, sigma
if (igrad.eq.1) then
  write(201) s, xi, fxdd, zeta, fx
end if
if (probe) write(201) v, lv, lsv
close(201)
!
end program ref_dump
