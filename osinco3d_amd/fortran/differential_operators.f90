! Drop-in module `diffoper` (reference: src/differential_operators.f90).
!
! divergence, rotational and calculate_Q_criterion each run one fused HIP
! kernel (o3d_divergence, o3d_rotational, o3d_q_criterion) that computes all
! the first derivatives it needs from one read of the fields, with the
! derivative variants the reference's routines select.  Boundary types come
! from the reference's initialization module, whose schemes() the driver
! calls unchanged.
module diffoper
  use initialization
  use derivation
  use o3d_capi
  implicit none
  private :: div_x, rot_x, q_x

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
    call rot_x(rotx, roty, rotz, ux, uy, uz, dx, dy, dz, nx, ny, nz)
  end subroutine rotational

  subroutine rot_x(rotx, roty, rotz, ux, uy, uz, dx, dy, dz, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(kind=8), intent(in) :: ux(nx, ny, nz), uy(nx, ny, nz), uz(nx, ny, nz)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(out) :: rotx(nx, ny, nz), roty(nx, ny, nz), rotz(nx, ny, nz)
    integer(c_int) :: bc(3)
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    call o3d_check(o3d_rotational(rotx, roty, rotz, ux, uy, uz, dx, dy, dz, int(nx, c_int), &
         int(ny, c_int), int(nz, c_int), bc), 'rotational')
  end subroutine rot_x

  !> Q = -1/2 (ux_x^2 + uy_y^2 + uz_z^2) - ux_y uy_x - ux_z uz_x - uy_z uz_y
  !> (src/differential_operators.f90:79-108).
  subroutine calculate_Q_criterion(Q, ux, uy, uz, dx, dy, dz, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(kind=8), intent(in) :: ux(:,:,:), uy(:,:,:), uz(:,:,:)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(out) :: Q(:,:,:)
    call q_x(Q, ux, uy, uz, dx, dy, dz, nx, ny, nz)
  end subroutine calculate_Q_criterion

  subroutine q_x(Q, ux, uy, uz, dx, dy, dz, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(kind=8), intent(in) :: ux(nx, ny, nz), uy(nx, ny, nz), uz(nx, ny, nz)
    real(kind=8), intent(in) :: dx, dy, dz
    real(kind=8), intent(out) :: Q(nx, ny, nz)
    integer(c_int) :: bc(3)
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    call o3d_check(o3d_q_criterion(Q, ux, uy, uz, dx, dy, dz, int(nx, c_int), int(ny, c_int), &
         int(nz, c_int), bc), 'calculate_Q_criterion')
  end subroutine q_x

end module diffoper
