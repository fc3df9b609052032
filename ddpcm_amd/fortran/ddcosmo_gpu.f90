! Fortran host of the MI355X ddCOSMO solver.
!
! Module `ddcosmo` keeps the public names, argument lists and module data of
! the reference module (ddcosmo.f90:87-127, cosmo.f90:1, forces_dd.f90:1,
! mkrhs.f90:1, efld.f90:1, matvec.f90:94-323), so that a caller written for
! the reference -- its own main.f90 included -- compiles and runs unchanged.
! The bodies call libddx.so through ddx_cabi; all compute runs on the GPU.
! As in the reference, the module holds one system at a time; failures print
! and stop.
module ddcosmo
  use, intrinsic :: iso_c_binding
  use ddx_cabi
  implicit none

  ! numerical constants (ddcosmo.f90:87-89)
  integer, parameter :: ndiis = 25, iout = 6, nngmax = 100
  real*8, parameter :: zero = 0.d0, pt5 = 0.5d0, one = 1.d0, two = 2.d0, four = 4.d0
  real*8, parameter :: se = -1.0d0
  real*8 :: pi, sq2

  ! control parameters, set by the caller before ddinit
  integer :: iprint, nproc, lmax, ngrid, iconv, igrad
  real*8 :: eps, eta

  ! sizes and module data, filled by ddinit
  integer :: nsph, ncav, nylm
  integer, allocatable :: inl(:), nl(:)
  real*8, allocatable :: rsph(:), csph(:,:), ccav(:,:)
  real*8, allocatable :: w(:), grid(:,:), basis(:,:)
  real*8, allocatable :: facl(:), facs(:)
  real*8, allocatable :: fi(:,:), ui(:,:), zi(:,:,:)
  logical :: grad

  type(c_ptr), save, private :: handle = c_null_ptr
  logical, save, private :: mixed = .false.

contains

  subroutine ddx_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= DDX_OK) then
      write(iout, '(1x,a,a,i3,a,a)') what, ': libddx error ', rc, ': ', trim(ddx_error_message())
      stop
    end if
  end subroutine ddx_check

  ! ---------------------------------------------------------------- setup
  ! ddinit (ddcosmo.f90:159): tables, neighbours, fi/ui/zi and the cavity on
  ! the GPU; the module arrays are host copies for the caller.
  subroutine ddinit(n, x, y, z, rvdw)
    integer, intent(in) :: n
    real*8, dimension(n), intent(in) :: x, y, z, rvdw
    integer(c_int) :: rc, ns, nc, ny, ng, lm, nnl
    integer :: istatus
    character(len=16) :: prec_env
    if (nproc == 0) nproc = 1
    pi = four*atan(one)
    sq2 = sqrt(two)
    if (c_associated(handle)) call memfree
    rc = ddx_init(handle, int(n, c_int), x, y, z, rvdw, int(iprint, c_int), int(lmax, c_int), &
                  int(ngrid, c_int), int(iconv, c_int), int(igrad, c_int), eps, eta)
    call ddx_check(rc, 'ddinit')
    ! DDX_PRECISION=mixed in the environment selects the mixed precision
    ! (FP32 operator rounds, FP64 residual); the default is the reference's FP64
    call get_environment_variable('DDX_PRECISION', prec_env)
    mixed = trim(prec_env) == 'mixed'
    if (mixed) call ddx_check(ddx_set_precision(handle, DDX_PREC_MIXED), 'ddinit')
    rc = ddx_get_sizes(handle, ns, nc, ny, ng, lm)
    nsph = ns ; ncav = nc ; nylm = ny ; ngrid = ng
    grad = igrad /= 0
    call header
    allocate(rsph(nsph), csph(3,nsph), ccav(3,ncav), w(ngrid), grid(3,ngrid), basis(nylm,ngrid), &
             facl(nylm), facs(nylm), fi(ngrid,nsph), ui(ngrid,nsph), inl(nsph+1), stat=istatus)
    if (istatus /= 0) then
      write(*,*) 'ddinit : allocation failed !'
      stop
    end if
    call ddx_check(ddx_get_array(handle, 'rsph'//c_null_char, rsph), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'csph'//c_null_char, csph), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'ccav'//c_null_char, ccav), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'w'//c_null_char, w), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'grid'//c_null_char, grid), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'basis'//c_null_char, basis), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'facl'//c_null_char, facl), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'facs'//c_null_char, facs), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'fi'//c_null_char, fi), 'ddinit')
    call ddx_check(ddx_get_array(handle, 'ui'//c_null_char, ui), 'ddinit')
    if (grad) then
      allocate(zi(3,ngrid,nsph))
      call ddx_check(ddx_get_array(handle, 'zi'//c_null_char, zi), 'ddinit')
    end if
    ! the neighbour list: inl first, then nl of inl(nsph+1)-1 entries
    allocate(nl(max(1, nsph*nngmax)))
    call ddx_check(ddx_get_neighbours(handle, inl, nl), 'ddinit')
    nnl = inl(nsph+1) - 1
    nl = nl(1:max(1,nnl))
    if (iprint >= 4) call dump_setup
  end subroutine ddinit

  ! the tables and the geometry that ddinit prints at iprint >= 4
  ! (ddcosmo.f90:274-280, 329-338, 413-416, 472-478), in its order and
  ! formats; the table labelled facl holds facs there, and so here
  subroutine dump_setup
    integer :: isph, ii
    call prtsph('facs', 1, 0, facs)
    call prtsph('facl', 1, 0, facs)
    call prtsph('basis', ngrid, 0, basis)
    call ptcart('grid', 3, 0, grid)
    call ptcart('weights', 1, 0, w)
    write(iout,*) '   inl:'
    write(iout,'(10i6)') inl(1:nsph+1)
    write(iout,*)
    do isph = 1, nsph
      write(iout,'(t3,"neighbours of sphere ",i6)') isph
      write(iout,'(t5,12i6)') nl(inl(isph):inl(isph+1)-1)
    end do
    write(iout,*)
    call ptcart('fi', nsph, 0, fi)
    call ptcart('ui', nsph, 0, ui)
    write(iout,*) '   external cavity points:'
    do ii = 1, ncav
      write(iout,'(t3,i8,3f14.6)') ii, ccav(:,ii)
    end do
    write(iout,*)
  end subroutine dump_setup

  ! memfree (ddcosmo.f90:491)
  subroutine memfree
    integer(c_int) :: rc
    if (c_associated(handle)) then
      rc = ddx_free(handle)
      handle = c_null_ptr
    end if
    if (allocated(rsph)) deallocate(rsph)
    if (allocated(csph)) deallocate(csph)
    if (allocated(ccav)) deallocate(ccav)
    if (allocated(w)) deallocate(w)
    if (allocated(grid)) deallocate(grid)
    if (allocated(basis)) deallocate(basis)
    if (allocated(inl)) deallocate(inl)
    if (allocated(nl)) deallocate(nl)
    if (allocated(facl)) deallocate(facl)
    if (allocated(facs)) deallocate(facs)
    if (allocated(ui)) deallocate(ui)
    if (allocated(fi)) deallocate(fi)
    if (allocated(zi)) deallocate(zi)
  end subroutine memfree

  ! ---------------------------------------------------------------- solver
  ! cosmo (cosmo.f90:1): L X = g (star false) or L* S = psi (star true);
  ! guess D^-1 rhs, Jacobi/DIIS to 10^-iconv, energy for the direct solve.
  subroutine cosmo(star, cart, phi, glm, psi, sigma, esolv)
    logical, intent(in) :: star, cart
    real*8, dimension(*), intent(in) :: phi
    real*8, dimension(nylm,nsph), intent(in) :: glm, psi
    real*8, dimension(nylm,nsph), intent(inout) :: sigma
    real*8, intent(inout) :: esolv
    integer(c_int) :: rc, istar, icart
    integer :: c1, c2, cr
    real*8 :: es
    real*8, allocatable :: rhs(:,:)
    call system_clock(count_rate=cr)
    call system_clock(count=c1)
    istar = merge(1, 0, star)
    icart = merge(1, 0, cart)
    es = zero
    rc = ddx_cosmo(handle, istar, icart, phi, glm, psi, sigma, es)
    ! the right-hand side, before the iterations (cosmo.f90:131,169)
    if (iprint >= 4 .and. (rc == 0 .or. rc == DDX_ERR_NOCONV)) then
      if (star) then
        call prtsph('rhs of the ddCOSMO adjoint equation', nsph, 0, psi)
      else
        allocate(rhs(nylm,nsph))
        call ddx_check(ddx_get_array(handle, 'rhs'//c_null_char, rhs), 'cosmo')
        call prtsph('rhs of the ddCOSMO equation', nsph, 0, rhs)
        deallocate(rhs)
      end if
    end if
    call print_history(istar)
    if (rc == DDX_ERR_NOCONV) then
      if (star) then
        write(iout, '(a)') ' adjoint ddCOSMO did not converge! Aborting...'
      else
        write(iout, '(a)') ' ddCOSMO did not converge! Aborting...'
      end if
      stop
    end if
    call ddx_check(rc, 'cosmo')
    if (.not. star) esolv = es
    call system_clock(count=c2)
    if (iprint > 0) then
      write(iout, *)
      if (star) then
        write(iout, '(a,f8.3,a)') ' solution time of ddCOSMO adjoint equations L^* \sigma = ', &
          dble(c2-c1)/dble(cr), ' secs.'
      else
        write(iout, '(a,f8.3,a)') ' solution time of ddCOSMO direct equations L \sigma = ', &
          dble(c2-c1)/dble(cr), ' secs.'
      end if
      write(iout, *)
    end if
  end subroutine cosmo

  ! the per-iteration increment norms, in the reference's format
  ! (jacobi_diis.f90:161-174, norm 4)
  subroutine print_history(istar)
    integer(c_int), intent(in) :: istar
    real(c_double) :: norms(256)
    integer(c_int) :: cnt, rc
    integer :: it
    if (iprint <= 0) return
    rc = ddx_get_history(handle, istar, 256_c_int, norms, cnt)
    do it = 1, min(int(cnt), 256)
      write(iout, '(t3,"iter=",i4," residual norm: ",d14.4)') it, norms(it)
    end do
  end subroutine print_history

  ! forces_dd (forces_dd.f90:1): fx(3,n) = 1/2 f(eps) <S, (L^(x) X - g^(x))>
  subroutine forces_dd(n, phi, sigma, s, fx)
    integer, intent(in) :: n
    real*8, dimension(ncav), intent(in) :: phi
    real*8, dimension(nylm,nsph), intent(in) :: sigma, s
    real*8, dimension(3,n), intent(inout) :: fx
    integer :: c1, c2, cr, isph
    real*8 :: fep
    real*8, allocatable :: xi(:,:)
    call system_clock(count_rate=cr)
    call system_clock(count=c1)
    call ddx_check(ddx_forces_dd(handle, phi, sigma, s, fx), 'forces_dd')
    call system_clock(count=c2)
    if (iprint >= 4) then
      ! xi = basis^T s, and the contributions before the factor 1/2 f(eps)
      ! (forces_dd.f90:116,141-147)
      allocate(xi(ngrid,nsph))
      call ddx_check(ddx_get_array(handle, 'xi'//c_null_char, xi), 'forces_dd')
      call ptcart('xi', nsph, 0, xi)
      deallocate(xi)
      fep = pt5*(eps - one)/eps
      write(iout, '(1x,"ddCOSMO-only contributions to the forces (atomic units):",/,1x," atom",15x,"x",15x,"y",15x,"z")')
      do isph = 1, nsph
        write(iout, '(1x,i5,3f16.8)') isph, fx(:,isph)/fep
      end do
    end if
    if (iprint > 0) write(iout, '(a,f8.3,a)') &
      ' the computation of the ddCOSMO part of the forces took ', dble(c2-c1)/dble(cr), ' seconds.'
  end subroutine forces_dd

  ! ddmkzeta (ddcosmo.f90:1787)
  subroutine ddmkzeta(s, zeta)
    real*8, dimension(nylm,nsph), intent(in) :: s
    real*8, dimension(ncav), intent(inout) :: zeta
    call ddx_check(ddx_mkzeta(handle, s, zeta), 'ddmkzeta')
  end subroutine ddmkzeta

  ! ---------------------------------------------------------------- solute
  ! mkrhs (mkrhs.f90:1): the potential at the nc points cc of the n point
  ! charges at (x, y, z), and psi; any coordinates, not only ddinit's
  subroutine mkrhs(n, charge, x, y, z, nc, cc, phi, ny, psi)
    integer, intent(in) :: n, nc, ny
    real*8, dimension(n), intent(in) :: charge, x, y, z
    real*8, dimension(3,nc), intent(in) :: cc
    real*8, dimension(nc), intent(inout) :: phi
    real*8, dimension(ny,n), intent(inout) :: psi
    call ddx_check(ddx_mkrhs_at(handle, int(n, c_int), charge, x, y, z, int(nc, c_int), cc, phi, &
                                int(ny, c_int), psi), 'mkrhs')
  end subroutine mkrhs

  ! efld (efld.f90:1): field of nsrc point sources at ntrg targets
  subroutine efld(nsrc, src, csrc, ntrg, ctrg, ef)
    integer, intent(in) :: nsrc, ntrg
    real*8, dimension(nsrc), intent(in) :: src
    real*8, dimension(3,nsrc), intent(in) :: csrc
    real*8, dimension(3,ntrg), intent(in) :: ctrg
    real*8, dimension(3,ntrg), intent(inout) :: ef
    call ddx_check(ddx_efld(handle, int(nsrc, c_int), src, csrc, int(ntrg, c_int), ctrg, ef), &
                   'efld')
  end subroutine efld

  ! ---------------------------------------------------------------- operators
  ! lx, lstarx, ldm1x, hnorm (matvec.f90:94-323), the callbacks of jacobi_diis
  subroutine lx(n, x, y)
    integer, intent(in) :: n
    real*8, dimension(nylm,nsph), intent(in) :: x
    real*8, dimension(nylm,nsph), intent(inout) :: y
    call ddx_check(ddx_lx(handle, x, y), 'lx')
  end subroutine lx

  subroutine lstarx(n, x, y)
    integer, intent(in) :: n
    real*8, dimension(nylm,nsph), intent(in) :: x
    real*8, dimension(nylm,nsph), intent(inout) :: y
    call ddx_check(ddx_lstarx(handle, x, y), 'lstarx')
  end subroutine lstarx

  subroutine ldm1x(n, x, y)
    integer, intent(in) :: n
    real*8, dimension(nylm,nsph), intent(in) :: x
    real*8, dimension(nylm,nsph), intent(inout) :: y
    call ddx_check(ddx_ldm1x(handle, x, y), 'ldm1x')
  end subroutine ldm1x

  real*8 function hnorm(n, x)
    integer, intent(in) :: n
    real*8, dimension(nylm,nsph), intent(in) :: x
    real(c_double) :: v
    call ddx_check(ddx_hnorm(handle, x, v), 'hnorm')
    hnorm = v
  end function hnorm

  ! jacobi_diis (jacobi_diis.f90:43-55): the reference's solver interface
  ! with caller-supplied matvec, dm1vec and (norm 4) u_norm.  With this
  ! module's own lx or lstarx, ldm1x and hnorm the whole solve runs on the
  ! GPU; any other operator is called back on the host while the iterate
  ! and the DIIS history stay on the GPU.  Residual lines as the
  ! reference's when lprint > 0.
  subroutine jacobi_diis(n, lprint, diis_max, norm, tol, rhs, x, n_iter, ok, matvec, dm1vec, &
                         u_norm)
    integer, intent(in) :: n, diis_max, norm, lprint
    real*8, intent(in) :: tol
    real*8, dimension(n), intent(in) :: rhs
    real*8, dimension(n), intent(inout) :: x
    integer, intent(inout) :: n_iter
    logical, intent(inout) :: ok
    external :: matvec, dm1vec
    real*8, optional :: u_norm
    external :: u_norm
    type(c_funptr) :: fm, fd, fn
    integer(c_int) :: it, iok, rc
    if (norm == 4 .and. .not. present(u_norm)) then
      write(iout,*) ' must provide a function norm(n,x) to evaluate the norm of the increment'
      stop
    end if
    ! this module's operators become the library's own callbacks, which the
    ! library recognises and runs on the device
    fm = c_funloc(matvec)
    if (c_associated(fm, c_funloc(lx))) fm = c_funloc(ddx_op_lx)
    if (c_associated(fm, c_funloc(lstarx))) fm = c_funloc(ddx_op_lstarx)
    fd = c_funloc(dm1vec)
    if (c_associated(fd, c_funloc(ldm1x))) fd = c_funloc(ddx_op_ldm1x)
    fn = c_null_funptr
    if (present(u_norm)) then
      fn = c_funloc(u_norm)
      if (c_associated(fn, c_funloc(hnorm))) fn = c_funloc(ddx_op_hnorm)
    end if
    flush(iout)
    it = int(n_iter, c_int)
    iok = 0
    rc = ddx_jacobi_diis_ops(handle, int(n, c_int), int(lprint, c_int), int(diis_max, c_int), &
                             int(norm, c_int), tol, rhs, x, it, iok, fm, fd, fn)
    if (rc /= DDX_OK .and. rc /= DDX_ERR_NOCONV) call ddx_check(rc, 'jacobi_diis')
    n_iter = it
    ok = iok /= 0
  end subroutine jacobi_diis

  ! sprod (ddcosmo.f90:531)
  real*8 function sprod(n, u, v)
    integer, intent(in) :: n
    real*8, dimension(n), intent(in) :: u, v
    integer :: i
    sprod = zero
    do i = 1, n
      sprod = sprod + u(i)*v(i)
    end do
  end function sprod

  ! ---------------------------------------------------------------- printing
  subroutine header
    if (iprint > 0) then
      write(iout, '(/,a,/)') ' ddCOSMO on AMD Instinct MI355X: domain-decomposition COSMO solvation.'
      write(iout, '(a)') ' Parameters:'
      write(iout, '(a,8x,i8)') '   number of grid points:                  ', ngrid
      write(iout, '(a,8x,i8)') '   number of spheres:                      ', nsph
      write(iout, '(a,8x,i8)') '   lmax for the spherical harmonics basis: ', lmax
      write(iout, '(a,8x,d8.1)') '   convergence threshold:                  ', 10.0d0**(-iconv)
      write(iout, '(a,8x,f8.3)') '   regularization parameter (eta):         ', eta
      write(iout, '(a,8x,f8.4,/)') '   dielectric constant:                    ', eps
      if (mixed) write(iout, '(a,/)') &
        ' Mixed precision (DDX_PRECISION=mixed): FP32 operator rounds, FP64 residual.'
      if (igrad == 1) write(iout, '(a,//)') ' Compute forces.'
    end if
  end subroutine header

  ! prtsph (ddcosmo.f90:724): an (nylm, ncol) array, or one column
  subroutine prtsph(label, ncol, icol, x)
    character(len=*), intent(in) :: label
    integer, intent(in) :: ncol, icol
    real*8, dimension(nylm,ncol), intent(in) :: x
    integer :: l, m, ind, c0, c1, j
    if (ncol == 1) then
      write(iout, '(3x,a,1x,"(column ",i4,")")') label, icol
      do l = 0, lmax
        ind = l*l + l + 1
        do m = -l, l
          write(iout, '(1x,i3,i4,f14.8)') l, m, x(ind+m,1)
        end do
      end do
    else
      ! blocks of five columns, then the block of the remaining ncol mod 5,
      ! which the reference writes also when it is empty
      write(iout, '(3x,a)') label
      do c0 = 1, ncol - mod(ncol, 5) + 1, 5
        c1 = min(ncol, c0 + 4)
        write(iout, '(8x,5i14)') (j, j = c0, c1)
        do l = 0, lmax
          ind = l*l + l + 1
          do m = -l, l
            write(iout, '(1x,i3,i4,5f14.8)') l, m, x(ind+m,c0:c1)
          end do
        end do
      end do
    end if
  end subroutine prtsph

  ! ptcart (ddcosmo.f90:667): an (ngrid, ncol) array, or one column
  subroutine ptcart(label, ncol, icol, x)
    character(len=*), intent(in) :: label
    integer, intent(in) :: ncol, icol
    real*8, dimension(ngrid,ncol), intent(in) :: x
    integer :: ig, c0, c1, j
    if (ncol == 1) then
      write(iout, '(3x,a,1x,"(column ",i4,")")') label, icol
      do ig = 1, ngrid
        write(iout, '(1x,i5,f14.8)') ig, x(ig,1)
      end do
    else
      write(iout, '(3x,a)') label
      do c0 = 1, ncol - mod(ncol, 5) + 1, 5
        c1 = min(ncol, c0 + 4)
        write(iout, '(6x,5i14)') (j, j = c0, c1)
        do ig = 1, ngrid
          write(iout, '(1x,i5,5f14.8)') ig, x(ig,c0:c1)
        end do
      end do
    end if
  end subroutine ptcart

end module ddcosmo
