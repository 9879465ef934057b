! Drop-in module `integration` (reference: src/integration.f90).
!
! predict_velocity, correct_pression and correct_velocity keep the
! reference's names and dummy-argument lists (src/integration.f90:14-16,
! 199-200, 257-258) and run the fused HIP kernels of libo3d.so:
!   predict_velocity  -> o3d_predict_velocity (9 D1 + 9 D2 + nu_t + AB update)
!   correct_pression  -> o3d_correct_pression (divergence/dt + red-black SOR), or
!                        o3d_divergence + o3d_poisson_multigrid with multigrid = 1
!   correct_velocity  -> o3d_correct_velocity (grad p, update, NaN/max guard)
! The run's fields (velocity, predicted velocity, history, pressure, nu_t)
! are registered with the library's device registry on first use, so they
! stay on the device between calls; the ones host code reads (velocity,
! pressure, nu_t) are copied back after every call that writes them, so the
! caller's arrays hold what the reference leaves in them.  Other host arrays
! are staged through device memory per call.  The passive scalar transeq runs o3d_transeq
! (derivatives, update, conservative clipping in one kernel chain).
module integration
  use functions
  use derivation
  use initialization
  use diffoper
  use les_turbulence
  use IOfunctions, only: print_in_outflow_rate, write_velocity_diverged
  use o3d_capi
  implicit none
  private :: predict_x, pression_x, pression_mg, correct_x

contains

  subroutine predict_velocity(ux_pred, uy_pred, uz_pred, ux, uy, uz, &
       fux, fuy, fuz, re, adt, bdt, cdt, itime, itscheme, &
       dx, dy, dz, nx, ny, nz, iles, cs, delta, nu_t)
    integer, intent(in) :: itscheme, itime
    integer, intent(in) :: nx, ny, nz, iles
    real(kind=8), intent(in) :: re, cs, delta
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), dimension(3), intent(in) :: adt, bdt, cdt
    real(kind=8), dimension(:,:,:), intent(inout), target :: ux, uy, uz
    real(kind=8), dimension(:,:,:), intent(inout), target :: nu_t
    real(kind=8), dimension(:,:,:,:), intent(inout), target :: fux, fuy, fuz
    real(kind=8), dimension(:,:,:), intent(out), target :: ux_pred, uy_pred, uz_pred

    print *, "* Predict velocity"
    if (itscheme == 1 .or. itime == 1) then
       print *, " Euler"
    elseif (itscheme == 2 .or. itime == 2) then
       print *, " Adams-Bashforth 2"
    elseif (itscheme == 3) then
       print *, " Adams-Bashforth 3"
    else
       print *, 'itscheme:', itscheme, ' unrecognized'
       stop
    end if
    if (iles == 1) print *, " LES model calculation"
    ! The run's fields stay on the device between calls: the velocity (read
    ! by the driver's host diagnostics every step, so mirrored back), the
    ! predicted velocity and the history (device only), nu_t (mirrored).
    call o3d_keep3(ux, 1)
    call o3d_keep3(uy, 1)
    call o3d_keep3(uz, 1)
    call o3d_keep3(ux_pred, 0)
    call o3d_keep3(uy_pred, 0)
    call o3d_keep3(uz_pred, 0)
    call o3d_keep4(fux, 0)
    call o3d_keep4(fuy, 0)
    call o3d_keep4(fuz, 0)
    if (iles == 1) call o3d_keep3(nu_t, 1)
    call predict_x(ux_pred, uy_pred, uz_pred, ux, uy, uz, fux, fuy, fuz, re, adt, bdt, &
         cdt, itime, itscheme, dx, dy, dz, nx, ny, nz, iles, cs, delta, nu_t)
    print *, ""
  end subroutine predict_velocity

  subroutine predict_x(ux_pred, uy_pred, uz_pred, ux, uy, uz, fux, fuy, fuz, re, adt, bdt, &
       cdt, itime, itscheme, dx, dy, dz, nx, ny, nz, iles, cs, delta, nu_t)
    integer, intent(in) :: itscheme, itime, nx, ny, nz, iles
    real(kind=8), intent(in) :: re, cs, delta, dx, dy, dz
    real(kind=8), intent(in) :: adt(3), bdt(3), cdt(3)
    real(kind=8), intent(in) :: ux(nx, ny, nz), uy(nx, ny, nz), uz(nx, ny, nz)
    real(kind=8), intent(inout) :: nu_t(nx, ny, nz)
    real(kind=8), intent(inout) :: fux(nx, ny, nz, 3), fuy(nx, ny, nz, 3), fuz(nx, ny, nz, 3)
    real(kind=8), intent(out) :: ux_pred(nx, ny, nz), uy_pred(nx, ny, nz), uz_pred(nx, ny, nz)
    integer(c_int) :: bc(3)
    real(kind=8) :: stats(6)
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    call o3d_check(o3d_predict_velocity(ux_pred, uy_pred, uz_pred, ux, uy, uz, fux, fuy, &
         fuz, re, adt, bdt, cdt, int(itime, c_int), int(itscheme, c_int), dx, dy, dz, &
         int(nx, c_int), int(ny, c_int), int(nz, c_int), int(iles, c_int), cs, delta, nu_t, &
         bc), 'predict_velocity')
    if (iles == 1) then
       stats = function_stats(nu_t, nx, ny, nz)
       call print_nu_t_statistics(stats)
    end if
  end subroutine predict_x

  subroutine correct_pression(pp, ux_pred, uy_pred, uz_pred, dx, dy, dz, &
       nx, ny, nz, dt, omega, eps, kmax, idyn, multigrid)
    real(kind=8), intent(inout), target :: pp(:,:,:)
    real(kind=8), intent(inout) :: omega
    real(kind=8), intent(in) :: ux_pred(:,:,:), uy_pred(:,:,:), uz_pred(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz, dt, eps
    integer, intent(in) :: nx, ny, nz, kmax, idyn, multigrid

    print *, "* Correction pression"
    call o3d_keep3(pp, 1)  ! the driver writes the pressure out on I/O steps
    if (multigrid == 1) then
       ! src/integration.f90:242-244: npre = 5, npost = 4, tol = eps, kmax as
       ! the level count (here also the cycle limit).
       call pression_mg(pp, ux_pred, uy_pred, uz_pred, dx, dy, dz, nx, ny, nz, dt, eps, kmax)
    else
       call pression_x(pp, ux_pred, uy_pred, uz_pred, dx, dy, dz, nx, ny, nz, dt, omega, eps, &
            kmax, idyn)
    end if
    print *, ""
  end subroutine correct_pression

  subroutine pression_mg(pp, ux_pred, uy_pred, uz_pred, dx, dy, dz, nx, ny, nz, dt, eps, kmax)
    integer, intent(in) :: nx, ny, nz, kmax
    real(kind=8), intent(inout) :: pp(nx, ny, nz)
    real(kind=8), intent(in) :: ux_pred(nx, ny, nz), uy_pred(nx, ny, nz), uz_pred(nx, ny, nz)
    real(kind=8), intent(in) :: dx, dy, dz, dt, eps
    real(kind=8), allocatable :: rhs(:,:,:)
    integer(c_int) :: bc(3), solver, cycles, reason
    real(c_double) :: stats(6), resmax
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    if (bc(1) == 0 .and. bc(2) == 0) then
       solver = 0
    else if (bc(1) == 0 .and. bc(2) == 1) then
       solver = 1
    else if (bc(1) == 1 .and. bc(2) == 0) then
       solver = -1
    else
       solver = 2
    end if
    allocate(rhs(nx, ny, nz))
    call o3d_check(o3d_divergence(rhs, ux_pred, uy_pred, uz_pred, dx, dy, dz, int(nx, c_int), &
         int(ny, c_int), int(nz, c_int), bc, 1_c_int, dt, stats), 'correct_pression')
    call o3d_check(o3d_poisson_multigrid(pp, rhs, dx, dy, dz, int(nx, c_int), int(ny, c_int), &
         int(nz, c_int), solver, int(kmax, c_int), 5_c_int, 4_c_int, eps, int(kmax, c_int), &
         cycles, resmax, reason), 'solve_poisson_multigrid')
    deallocate(rhs)
    write(*, '(A20,I4,A1,E13.6,A1,I2)') ' Multigrid cycles : ', cycles, ',', resmax, ',', reason
  end subroutine pression_mg

  subroutine pression_x(pp, ux_pred, uy_pred, uz_pred, dx, dy, dz, nx, ny, nz, dt, omega, &
       eps, kmax, idyn)
    integer, intent(in) :: nx, ny, nz, kmax, idyn
    real(kind=8), intent(inout) :: pp(nx, ny, nz), omega
    real(kind=8), intent(in) :: ux_pred(nx, ny, nz), uy_pred(nx, ny, nz), uz_pred(nx, ny, nz)
    real(kind=8), intent(in) :: dx, dy, dz, dt, eps
    integer(c_int) :: bc(3), it, reason
    real(c_double) :: dmax
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    if (bc(1) == 0 .and. bc(2) == 0) then
       print *, "* Poisson solver 0000 start"
    else if (bc(1) == 0 .and. bc(2) == 1) then
       print *, "* Poisson solver 0011 start"
    else
       print *, "* Poisson solver 111111 start"
    end if
    call o3d_check(o3d_correct_pression(pp, ux_pred, uy_pred, uz_pred, dx, dy, dz, &
         int(nx, c_int), int(ny, c_int), int(nz, c_int), dt, omega, eps, int(kmax, c_int), &
         int(idyn, c_int), bc, it, dmax, reason), 'correct_pression')
    if (reason == 1) write(*, '(A29,E10.3)') " No converged at less than: ", eps
    write(*, '(A19,I4,A1,E13.6)') ' Exit iteration : ', it, ',', dmax
    if (idyn == 1) write(*,'(A23, F5.3)') ' With dynamic omega = ', omega
  end subroutine pression_x

  subroutine correct_velocity(ux, uy, uz, ux_pred, uy_pred, uz_pred, &
       pp, dt, dx, dy, dz, nx, ny, nz)
    real(kind=8), intent(inout), target :: ux(:,:,:), uy(:,:,:), uz(:,:,:)
    real(kind=8), intent(in) :: ux_pred(:,:,:), uy_pred(:,:,:), uz_pred(:,:,:)
    real(kind=8), intent(in) :: pp(:,:,:)
    real(kind=8), intent(in) :: dt, dx, dy, dz
    integer, intent(in) :: nx, ny, nz
    integer :: rc

    print *, "* Correct velocity"
    call o3d_keep3(ux, 1)
    call o3d_keep3(uy, 1)
    call o3d_keep3(uz, 1)
    call correct_x(ux, uy, uz, ux_pred, uy_pred, uz_pred, pp, dt, dx, dy, dz, nx, ny, nz, rc)
    if (rc == 1) then
       call write_velocity_diverged()
       stop
    end if
    print*, ""
  end subroutine correct_velocity

  subroutine correct_x(ux, uy, uz, ux_pred, uy_pred, uz_pred, pp, dt, dx, dy, dz, nx, ny, &
       nz, rc)
    integer, intent(in) :: nx, ny, nz
    real(kind=8), intent(inout) :: ux(nx, ny, nz), uy(nx, ny, nz), uz(nx, ny, nz)
    real(kind=8), intent(in) :: ux_pred(nx, ny, nz), uy_pred(nx, ny, nz), uz_pred(nx, ny, nz)
    real(kind=8), intent(in) :: pp(nx, ny, nz), dt, dx, dy, dz
    integer, intent(out) :: rc
    integer(c_int) :: bc(3), r
    real(c_double) :: vs(12)
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    r = o3d_correct_velocity(ux, uy, uz, ux_pred, uy_pred, uz_pred, pp, dt, dx, dy, dz, &
         int(nx, c_int), int(ny, c_int), int(nz, c_int), bc, vs)
    call o3d_check(r, 'correct_velocity')
    rc = r
  end subroutine correct_x

  !> Passive scalar transport (src/integration.f90:332-468) -> o3d_transeq.
  subroutine transeq(phi, ux, uy, uz, src, fphi, re, sc, adt, bdt, cdt, &
       itime, itscheme, dx, dy, dz, nx, ny, nz, iles, nu_t)
    integer, intent(in) :: itime, itscheme, nx, ny, nz, iles
    real(kind=8), intent(in) :: re, sc, dx, dy, dz
    real(kind=8), dimension(3), intent(in) :: adt, bdt, cdt
    real(kind=8), dimension(nx,ny,nz), intent(in) :: ux, uy, uz, src, nu_t
    real(kind=8), dimension(nx,ny,nz), intent(inout) :: phi
    real(kind=8), dimension(nx,ny,nz,3), intent(inout) :: fphi
    integer(c_int) :: bc(3), rc
    print *, "* Transport equation"
    if (itscheme == 1 .or. itime == 1) then
       print *, " Euler"
    elseif (itscheme == 2 .or. itime == 2) then
       print *, " Adams-Bashforth 2"
    elseif (itscheme == 3) then
       print *, " Adams-Bashforth 3"
    end if
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    rc = o3d_transeq(phi, ux, uy, uz, src, fphi, re, sc, adt, bdt, cdt, int(itime, c_int), &
         int(itscheme, c_int), dx, dy, dz, int(nx, c_int), int(ny, c_int), int(nz, c_int), &
         int(iles, c_int), nu_t, bc)
    if (rc == -2) then
       print *, 'itscheme:', itscheme, ' unrecognized'
       stop
    end if
    call o3d_check(rc, 'transeq')
    print *, ""
  end subroutine transeq

end module integration
