! Drop-in module `diffoper` (reference: src/differential_operators.f90).
!
! divergence runs the fused HIP divergence kernel (o3d_divergence); the
! diagnostic rotational and calculate_Q_criterion (called every nfre steps by
! the driver) combine derivatives from the drop-in `derivation` pointers on
! the host, as the reference does.  Boundary types come from the reference's
! initialization module, whose schemes() the driver calls unchanged.
module diffoper
  use initialization
  use derivation
  use o3d_capi
  implicit none
  private :: div_x

contains

  subroutine divergence(divf, fx, fy, fz, dx, dy, dz, nx, ny, nz, odd)
    integer, intent(in) :: nx, ny, nz, odd
    real(kind=8), intent(in) :: fx(:,:,:), fy(:,:,:), fz(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(out) :: divf(:,:,:)
    call div_x(divf, fx, fy, fz, dx, dy, dz, nx, ny, nz, odd)
  end subroutine divergence

  subroutine div_x(divf, fx, fy, fz, dx, dy, dz, nx, ny, nz, odd)
    integer, intent(in) :: nx, ny, nz, odd
    real(kind=8), intent(in) :: fx(nx, ny, nz), fy(nx, ny, nz), fz(nx, ny, nz)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(out) :: divf(nx, ny, nz)
    integer(c_int) :: bc(3)
    real(c_double) :: st(6)
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    call o3d_check(o3d_divergence(divf, fx, fy, fz, dx, dy, dz, int(nx, c_int), &
         int(ny, c_int), int(nz, c_int), bc, int(odd, c_int), 1.d0, st), 'divergence')
  end subroutine div_x

  !> Curl of (ux, uy, uz) with the even ("p") first derivatives
  !> (src/differential_operators.f90:40-77).
  subroutine rotational(rotx, roty, rotz, ux, uy, uz, dx, dy, dz, nx, ny, nz)
    real(kind=8), intent(in) :: ux(:,:,:), uy(:,:,:), uz(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(out) :: rotx(:,:,:), roty(:,:,:), rotz(:,:,:)
    integer, intent(in) :: nx, ny, nz
    real(kind=8), allocatable :: a(:,:,:), b(:,:,:)
    allocate(a(nx, ny, nz), b(nx, ny, nz))
    call deryp(a, uz, dy)
    call derzp(b, uy, dz)
    rotx = a - b
    call derzp(a, ux, dz)
    call derxp(b, uz, dx)
    roty = a - b
    call derxp(a, uy, dx)
    call deryp(b, ux, dy)
    rotz = a - b
    deallocate(a, b)
  end subroutine rotational

  !> Q = -1/2 (ux_x^2 + uy_y^2 + uz_z^2) - ux_y uy_x - ux_z uz_x - uy_z uz_y
  !> (src/differential_operators.f90:79-108).
  subroutine calculate_Q_criterion(Q, ux, uy, uz, dx, dy, dz, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(kind=8), intent(in) :: ux(:,:,:), uy(:,:,:), uz(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(out) :: Q(:,:,:)
    real(kind=8), allocatable :: g(:,:,:,:)
    allocate(g(nx, ny, nz, 9))
    call derxi(g(:,:,:,1), ux, dx)
    call deryi(g(:,:,:,2), uy, dy)
    call derzi(g(:,:,:,3), uz, dz)
    call deryp(g(:,:,:,4), ux, dy)
    call derzp(g(:,:,:,5), ux, dz)
    call derxp(g(:,:,:,6), uy, dx)
    call derzp(g(:,:,:,7), uy, dz)
    call derxp(g(:,:,:,8), uz, dx)
    call deryp(g(:,:,:,9), uz, dy)
    Q = - 0.5d0 * (g(:,:,:,1) ** 2 + g(:,:,:,2) ** 2 + g(:,:,:,3) ** 2) - &
         g(:,:,:,4) * g(:,:,:,6) - g(:,:,:,5) * g(:,:,:,8) - g(:,:,:,7) * g(:,:,:,9)
    deallocate(g)
  end subroutine calculate_Q_criterion

end module diffoper
