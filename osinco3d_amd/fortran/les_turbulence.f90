! Drop-in module `les_turbulence` (reference: src/les_turbulence.f90).
!
! calculate_nu_t runs the Smagorinsky viscosity on the GPU (o3d_nu_t: the
! nine first derivatives, |S| and (cs*delta)^2 |S| in one fused kernel) and
! prints the reference's nu_t statistics.  calculate_tau_ij and
! calculate_dtau_ij_dxj have no caller in the reference and are not provided.
module les_turbulence
  use IOfunctions
  use derivation
  use initialization
  use functions
  use o3d_capi
  implicit none
  private :: nu_t_x

contains

  subroutine calculate_nu_t(nu_t, ux, uy, uz, dx, dy, dz, cs, delta)
    real(kind=8), dimension(:,:,:), intent(in) :: ux, uy, uz
    real(kind=8), intent(in) :: dx, dy, dz, cs, delta
    real(kind=8), dimension(:,:,:), intent(out) :: nu_t
    real(kind=8) :: stats(6)
    call nu_t_x(nu_t, ux, uy, uz, dx, dy, dz, cs, delta, size(ux, 1), size(ux, 2), &
         size(ux, 3), stats)
    call print_nu_t_statistics(stats)
  end subroutine calculate_nu_t

  subroutine nu_t_x(nu_t, ux, uy, uz, dx, dy, dz, cs, delta, n1, n2, n3, stats)
    integer, intent(in) :: n1, n2, n3
    real(kind=8), intent(in) :: ux(n1, n2, n3), uy(n1, n2, n3), uz(n1, n2, n3)
    real(kind=8), intent(in) :: dx, dy, dz, cs, delta
    real(kind=8), intent(out) :: nu_t(n1, n2, n3), stats(6)
    integer(c_int) :: bc(3)
    bc = [int(nbcx1, c_int), int(nbcy1, c_int), int(nbcz1, c_int)]
    call o3d_check(o3d_nu_t(nu_t, ux, uy, uz, dx, dy, dz, cs, delta, int(n1, c_int), &
         int(n2, c_int), int(n3, c_int), bc, stats), 'calculate_nu_t')
  end subroutine nu_t_x

end module les_turbulence
