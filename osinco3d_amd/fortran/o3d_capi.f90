! ISO_C_BINDING interfaces of libo3d.so (include/o3d.h), used by the drop-in
! modules derivation, poisson, diffoper, les_turbulence and integration.
! Arrays are passed as assumed-size real(c_double) dummies, so an explicit-
! shape or contiguous actual argument is handed over in place.
module o3d_capi
  use iso_c_binding
  implicit none

  interface
     integer(c_int) function o3d_der(which, df, f, nx, ny, nz, d) bind(C, name="o3d_der")
       import :: c_int, c_double
       integer(c_int), value :: which, nx, ny, nz
       real(c_double) :: df(*)
       real(c_double), intent(in) :: f(*)
       real(c_double), value :: d
     end function o3d_der

     integer(c_int) function o3d_divergence(divf, fx, fy, fz, dx, dy, dz, nx, ny, nz, bc, &
          odd, divisor, stats6) bind(C, name="o3d_divergence")
       import :: c_int, c_double
       real(c_double) :: divf(*)
       real(c_double), intent(in) :: fx(*), fy(*), fz(*)
       real(c_double), value :: dx, dy, dz, divisor
       integer(c_int), value :: nx, ny, nz, odd
       integer(c_int), intent(in) :: bc(3)
       real(c_double) :: stats6(6)
     end function o3d_divergence

     integer(c_int) function o3d_transeq(phi, ux, uy, uz, src, fphi, re, sc, adt, bdt, cdt, &
          itime, itscheme, dx, dy, dz, nx, ny, nz, iles, nu_t, bc) bind(C, name="o3d_transeq")
       import :: c_int, c_double
       real(c_double) :: phi(*), fphi(*)
       real(c_double), intent(in) :: ux(*), uy(*), uz(*), src(*), nu_t(*)
       real(c_double), value :: re, sc, dx, dy, dz
       real(c_double), intent(in) :: adt(3), bdt(3), cdt(3)
       integer(c_int), value :: itime, itscheme, nx, ny, nz, iles
       integer(c_int), intent(in) :: bc(3)
     end function o3d_transeq

     integer(c_int) function o3d_rotational(rotx, roty, rotz, ux, uy, uz, dx, dy, dz, nx, ny, &
          nz, bc) bind(C, name="o3d_rotational")
       import :: c_int, c_double
       real(c_double) :: rotx(*), roty(*), rotz(*)
       real(c_double), intent(in) :: ux(*), uy(*), uz(*)
       real(c_double), value :: dx, dy, dz
       integer(c_int), value :: nx, ny, nz
       integer(c_int), intent(in) :: bc(3)
     end function o3d_rotational

     integer(c_int) function o3d_q_criterion(q, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc) &
          bind(C, name="o3d_q_criterion")
       import :: c_int, c_double
       real(c_double) :: q(*)
       real(c_double), intent(in) :: ux(*), uy(*), uz(*)
       real(c_double), value :: dx, dy, dz
       integer(c_int), value :: nx, ny, nz
       integer(c_int), intent(in) :: bc(3)
     end function o3d_q_criterion

     integer(c_int) function o3d_nu_t(nu_t, ux, uy, uz, dx, dy, dz, cs, delta, nx, ny, nz, &
          bc, stats6) bind(C, name="o3d_nu_t")
       import :: c_int, c_double
       real(c_double) :: nu_t(*)
       real(c_double), intent(in) :: ux(*), uy(*), uz(*)
       real(c_double), value :: dx, dy, dz, cs, delta
       integer(c_int), value :: nx, ny, nz
       integer(c_int), intent(in) :: bc(3)
       real(c_double) :: stats6(6)
     end function o3d_nu_t

     integer(c_int) function o3d_predict_velocity(ux_pred, uy_pred, uz_pred, ux, uy, uz, &
          fux, fuy, fuz, re, adt, bdt, cdt, itime, itscheme, dx, dy, dz, nx, ny, nz, iles, &
          cs, delta, nu_t, bc) bind(C, name="o3d_predict_velocity")
       import :: c_int, c_double
       real(c_double) :: ux_pred(*), uy_pred(*), uz_pred(*)
       real(c_double), intent(in) :: ux(*), uy(*), uz(*)
       real(c_double) :: fux(*), fuy(*), fuz(*), nu_t(*)
       real(c_double), value :: re, dx, dy, dz, cs, delta
       real(c_double), intent(in) :: adt(3), bdt(3), cdt(3)
       integer(c_int), value :: itime, itscheme, nx, ny, nz, iles
       integer(c_int), intent(in) :: bc(3)
     end function o3d_predict_velocity

     integer(c_int) function o3d_poisson(pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, &
          idyn, solver, exit_iter, dmax, reason) bind(C, name="o3d_poisson")
       import :: c_int, c_double
       real(c_double) :: pp(*)
       real(c_double), intent(in) :: rhs(*)
       real(c_double), value :: dx, dy, dz, eps
       integer(c_int), value :: nx, ny, nz, kmax, idyn, solver
       real(c_double) :: omega, dmax
       integer(c_int) :: exit_iter, reason
     end function o3d_poisson

     integer(c_int) function o3d_poisson_multigrid(pp, rhs, dx, dy, dz, nx, ny, nz, solver, &
          nlevels, npre, npost, tol, max_cycles, cycles, resmax, reason) &
          bind(C, name="o3d_poisson_multigrid")
       import :: c_int, c_double
       real(c_double) :: pp(*)
       real(c_double), intent(in) :: rhs(*)
       real(c_double), value :: dx, dy, dz, tol
       integer(c_int), value :: nx, ny, nz, solver, nlevels, npre, npost, max_cycles
       integer(c_int) :: cycles, reason
       real(c_double) :: resmax
     end function o3d_poisson_multigrid

     integer(c_int) function o3d_correct_pression(pp, ux_pred, uy_pred, uz_pred, dx, dy, dz, &
          nx, ny, nz, dt, omega, eps, kmax, idyn, bc, exit_iter, dmax, reason) &
          bind(C, name="o3d_correct_pression")
       import :: c_int, c_double
       real(c_double) :: pp(*)
       real(c_double), intent(in) :: ux_pred(*), uy_pred(*), uz_pred(*)
       real(c_double), value :: dx, dy, dz, dt, eps
       integer(c_int), value :: nx, ny, nz, kmax, idyn
       real(c_double) :: omega, dmax
       integer(c_int), intent(in) :: bc(3)
       integer(c_int) :: exit_iter, reason
     end function o3d_correct_pression

     integer(c_int) function o3d_correct_velocity(ux, uy, uz, ux_pred, uy_pred, uz_pred, pp, &
          dt, dx, dy, dz, nx, ny, nz, bc, vstats) bind(C, name="o3d_correct_velocity")
       import :: c_int, c_double
       real(c_double) :: ux(*), uy(*), uz(*)
       real(c_double), intent(in) :: ux_pred(*), uy_pred(*), uz_pred(*), pp(*)
       real(c_double), value :: dt, dx, dy, dz
       integer(c_int), value :: nx, ny, nz
       integer(c_int), intent(in) :: bc(3)
       real(c_double) :: vstats(12)
     end function o3d_correct_velocity

     type(c_ptr) function o3d_last_error() bind(C, name="o3d_last_error")
       import :: c_ptr
     end function o3d_last_error

     integer(c_int) function o3d_register(host, bytes, mirror) bind(C, name="o3d_register")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: host
       integer(c_size_t), value :: bytes
       integer(c_int), value :: mirror
     end function o3d_register

     integer(c_int) function o3d_transfer_stats(h2d, d2h, reset) &
          bind(C, name="o3d_transfer_stats")
       import :: c_int, c_double
       real(c_double) :: h2d, d2h
       integer(c_int), value :: reset
     end function o3d_transfer_stats
  end interface

contains

  ! Register a field the driver keeps for the whole run (a contiguous
  ! allocatable of src/initialization.f90:50-59) with the library's device
  ! registry: its device copy then lives across calls.  mirror = 1 when host
  ! code reads the field between steps (the library copies it back after
  ! each call that writes it).  A non-contiguous section is left alone (it
  ! is staged per call).
  subroutine o3d_keep3(a, mirror)
    real(c_double), dimension(:,:,:), intent(in), target :: a
    integer, intent(in) :: mirror
    if (.not. is_contiguous(a) .or. size(a) == 0) return
    call o3d_check(o3d_register(c_loc(a), int(8, c_size_t) * size(a, kind=c_size_t), &
         int(mirror, c_int)), 'o3d_register')
  end subroutine o3d_keep3

  subroutine o3d_keep4(a, mirror)
    real(c_double), dimension(:,:,:,:), intent(in), target :: a
    integer, intent(in) :: mirror
    if (.not. is_contiguous(a) .or. size(a) == 0) return
    call o3d_check(o3d_register(c_loc(a), int(8, c_size_t) * size(a, kind=c_size_t), &
         int(mirror, c_int)), 'o3d_register')
  end subroutine o3d_keep4

  ! Stop with the library's message after a runtime error (status < 0).
  subroutine o3d_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc >= 0) return
    call c_f_pointer(o3d_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    write(*, '(A,A,A,1024A1)') ' libo3d error in ', where, ': ', msg(1:n)
    stop
  end subroutine o3d_check

end module o3d_capi
