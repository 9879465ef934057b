! Drop-in module `poisson` (reference: src/poisson.f90).
!
! The three SOR solvers keep the reference's names and the poi_type interface
! (src/initialization.f90:93-102) that schemes() binds to its poisson_solver
! pointer.  Each runs libo3d.so's red-black SOR (o3d_poisson, include/o3d.h)
! with the reference's point update, neighbour rules, stop rules and dynamic
! omega, and prints the reference's exit lines.
module poisson
  use o3d_capi
  implicit none
  private :: run_sor, run_sor_x

contains

  subroutine run_sor(kind, label, pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
    integer, intent(in) :: kind
    character(len=*), intent(in) :: label
    real(kind=8), intent(inout) :: pp(:,:,:), omega
    real(kind=8), intent(in) :: rhs(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz, eps
    integer, intent(in) :: nx, ny, nz, kmax, idyn
    print *, "* Poisson solver " // label // " start"
    call run_sor_x(kind, pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
  end subroutine run_sor

  subroutine run_sor_x(kind, pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
    integer, intent(in) :: kind, nx, ny, nz, kmax, idyn
    real(kind=8), intent(inout) :: pp(nx, ny, nz), omega
    real(kind=8), intent(in) :: rhs(nx, ny, nz)
    real(kind=8), intent(in) :: dx, dy, dz, eps
    integer(c_int) :: it, reason
    real(c_double) :: dmax
    call o3d_check(o3d_poisson(pp, rhs, dx, dy, dz, int(nx, c_int), int(ny, c_int), &
         int(nz, c_int), omega, eps, int(kmax, c_int), int(idyn, c_int), int(kind, c_int), &
         it, dmax, reason), 'poisson')
    if (reason == 1) write(*, '(A29,E10.3)') " No converged at less than: ", eps
    if (kind == 2) then
       write(*, '(A18,I4,A1,E13.6)') ' Exit iteration: ', it, ',', dmax
    else
       write(*, '(A19,I4,A1,E13.6)') ' Exit iteration : ', it, ',', dmax
    end if
    if (idyn == 1) write(*,'(A23, F5.3)') ' With dynamic omega = ', omega
  end subroutine run_sor_x

  !> Periodic in x, y and z (src/poisson.f90:6-130).
  subroutine poisson_solver_0000(pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
    real(kind=8), intent(inout) :: pp(:,:,:), omega
    real(kind=8), intent(in) :: rhs(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(in) :: eps
    integer, intent(in) :: nx, ny, nz, kmax, idyn
    call run_sor(0, "0000", pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
  end subroutine poisson_solver_0000

  !> Periodic in x and z, free-slip in y (src/poisson.f90:132-255).
  subroutine poisson_solver_0011(pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
    real(kind=8), intent(inout) :: pp(:,:,:), omega
    real(kind=8), intent(in) :: rhs(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(in) :: eps
    integer, intent(in) :: nx, ny, nz, kmax, idyn
    call run_sor(1, "0011", pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
  end subroutine poisson_solver_0011

  !> Free-slip in x, y and z (src/poisson.f90:257-381).
  subroutine poisson_solver_111111(pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
    real(kind=8), intent(inout) :: pp(:,:,:), omega
    real(kind=8), intent(in) :: rhs(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(in) :: eps
    integer, intent(in) :: nx, ny, nz, kmax, idyn
    call run_sor(2, "111111", pp, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn)
  end subroutine poisson_solver_111111

end module poisson
