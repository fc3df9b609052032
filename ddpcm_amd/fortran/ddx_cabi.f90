! iso_c_binding interfaces of libddx.so (include/ddx.h).
module ddx_cabi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: DDX_OK = 0, DDX_ERR_ARG = 1, DDX_ERR_NOCONV = 3
  integer(c_int), parameter :: DDX_PREC_FP64 = 0, DDX_PREC_MIXED = 1

  interface
    function ddx_last_error() bind(C, name='ddx_last_error')
      import :: c_ptr
      type(c_ptr) :: ddx_last_error
    end function

    integer(c_int) function ddx_init(h, n, x, y, z, rvdw, iprint, lmax, ngrid, iconv, igrad, &
                                     eps, eta) bind(C, name='ddx_init')
      import :: c_ptr, c_int, c_double
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: n, iprint, lmax, ngrid, iconv, igrad
      real(c_double), intent(in) :: x(*), y(*), z(*), rvdw(*)
      real(c_double), value :: eps, eta
    end function

    integer(c_int) function ddx_free(h) bind(C, name='ddx_free')
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function

    integer(c_int) function ddx_set_precision(h, prec) bind(C, name='ddx_set_precision')
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: prec
    end function

    integer(c_int) function ddx_get_sizes(h, nsph, ncav, nylm, ngrid, lmax) &
        bind(C, name='ddx_get_sizes')
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nsph, ncav, nylm, ngrid, lmax
    end function

    integer(c_int) function ddx_get_array(h, name, dst) bind(C, name='ddx_get_array')
      import :: c_ptr, c_int, c_double, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: dst(*)
    end function

    integer(c_int) function ddx_get_neighbours(h, inl, nl) bind(C, name='ddx_get_neighbours')
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: inl(*), nl(*)
    end function

    integer(c_int) function ddx_mkrhs(h, charge, phi, psi) bind(C, name='ddx_mkrhs')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: charge(*)
      real(c_double), intent(out) :: phi(*), psi(*)
    end function

    integer(c_int) function ddx_cosmo(h, star, cart, phi, glm, psi, sigma, esolv) &
        bind(C, name='ddx_cosmo')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: star, cart
      real(c_double), intent(in) :: phi(*), glm(*), psi(*)
      real(c_double), intent(out) :: sigma(*), esolv
    end function

    integer(c_int) function ddx_forces_dd(h, phi, sigma, s, fx) bind(C, name='ddx_forces_dd')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: phi(*), sigma(*), s(*)
      real(c_double), intent(out) :: fx(*)
    end function

    integer(c_int) function ddx_mkzeta(h, s, zeta) bind(C, name='ddx_mkzeta')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: s(*)
      real(c_double), intent(out) :: zeta(*)
    end function

    integer(c_int) function ddx_efld(h, nsrc, src, csrc, ntrg, ctrg, ef) bind(C, name='ddx_efld')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nsrc, ntrg
      real(c_double), intent(in) :: src(*), csrc(*), ctrg(*)
      real(c_double), intent(out) :: ef(*)
    end function

    integer(c_int) function ddx_lx(h, x, y) bind(C, name='ddx_lx')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function

    integer(c_int) function ddx_lstarx(h, x, y) bind(C, name='ddx_lstarx')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function

    integer(c_int) function ddx_ldm1x(h, x, y) bind(C, name='ddx_ldm1x')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function

    integer(c_int) function ddx_hnorm(h, x, norm) bind(C, name='ddx_hnorm')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: norm
    end function

    integer(c_int) function ddx_jacobi_diis(h, star, diis_max, tol, rhs, x, n_iter, ok) &
        bind(C, name='ddx_jacobi_diis')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: star, diis_max
      real(c_double), value :: tol
      real(c_double), intent(in) :: rhs(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), intent(inout) :: n_iter
      integer(c_int), intent(out) :: ok
    end function

    integer(c_int) function ddx_jacobi_diis_ops(h, n, lprint, diis_max, norm, tol, rhs, x, &
                                                n_iter, ok, matvec, dm1vec, u_norm) &
        bind(C, name='ddx_jacobi_diis_ops')
      import :: c_ptr, c_int, c_double, c_funptr
      type(c_ptr), value :: h
      integer(c_int), value :: n, lprint, diis_max, norm
      real(c_double), value :: tol
      real(c_double), intent(in) :: rhs(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), intent(inout) :: n_iter
      integer(c_int), intent(out) :: ok
      type(c_funptr), value :: matvec, dm1vec, u_norm
    end function

    ! the handle's operators as callbacks of ddx_jacobi_diis_ops
    subroutine ddx_op_lx(n, x, y) bind(C, name='ddx_op_lx')
      import :: c_int, c_double
      integer(c_int), intent(in) :: n
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
    end subroutine
    subroutine ddx_op_lstarx(n, x, y) bind(C, name='ddx_op_lstarx')
      import :: c_int, c_double
      integer(c_int), intent(in) :: n
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
    end subroutine
    subroutine ddx_op_ldm1x(n, x, y) bind(C, name='ddx_op_ldm1x')
      import :: c_int, c_double
      integer(c_int), intent(in) :: n
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
    end subroutine
    real(c_double) function ddx_op_hnorm(n, x) bind(C, name='ddx_op_hnorm')
      import :: c_int, c_double
      integer(c_int), intent(in) :: n
      real(c_double), intent(in) :: x(*)
    end function

    integer(c_int) function ddx_mkrhs_at(h, n, charge, x, y, z, ncav, ccav, phi, nylm, psi) &
        bind(C, name='ddx_mkrhs_at')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n, ncav, nylm
      real(c_double), intent(in) :: charge(*), x(*), y(*), z(*), ccav(*)
      real(c_double), intent(out) :: phi(*), psi(*)
    end function

    integer(c_int) function ddx_get_history(h, star, cap, norms, count) &
        bind(C, name='ddx_get_history')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: star, cap
      real(c_double), intent(out) :: norms(*)
      integer(c_int), intent(out) :: count
    end function

    integer(c_int) function ddx_energy_forces(h, charge, esolv, fx) &
        bind(C, name='ddx_energy_forces')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: charge(*)
      real(c_double), intent(out) :: esolv, fx(*)
    end function
  end interface

contains

  ! the last error message of libddx as a Fortran string
  function ddx_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(ddx_last_error(), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function

end module ddx_cabi
