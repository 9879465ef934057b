! Test infrastructure only (the oracle, never the product path).
!
! bind(C) wrappers around the reference osinco3d Fortran modules, compiled from
! the reference sources in place by oracle/Makefile into oracle/_ref/.  They let
! the tests, the golden-vector generator and the CPU baseline call the
! reference's own routines on arrays handed over from Python (ctypes):
!   src/derivation.f90            (the 18 3D stencil routines, + 2D stubs)
!   src/differential_operators.f90:7-38     divergence
!   src/les_turbulence.f90:10-97            calculate_nu_t
!   src/integration.f90:14-330              predict_velocity, correct_pression,
!                                           correct_velocity
!   src/poisson.f90                         the three SOR solvers
!   src/functions.f90:27-63                 function_stats
!   src/utils.f90:243-375                   statistics_calc (writes outputs/stats.dat)
! Boundary conditions are bound exactly as the driver does it: set the
! initialization-module flags, then call schemes() (src/initialization.f90:226-304).
module o3d_ref_harness
  use iso_c_binding
  use initialization
  use derivation
  use poisson
  use diffoper
  use les_turbulence
  use integration
  use functions
  use utils
  implicit none
contains

  subroutine ref_set_bc(bx, by, bz, s2d) bind(C, name="ref_set_bc")
    integer(c_int), value :: bx, by, bz, s2d
    nbcx1 = bx; nbcxn = bx
    nbcy1 = by; nbcyn = by
    nbcz1 = bz; nbczn = bz
    sim2d = s2d
    poisson_solver => null()
    call schemes()
  end subroutine ref_set_bc

  ! which = axis*6 + (order-1)*3 + variant; variant 0 = _00, 1 = p_11, 2 = i_11
  subroutine ref_der(which, nx, ny, nz, f, df, h) bind(C, name="ref_der")
    integer(c_int), value :: which, nx, ny, nz
    real(c_double), intent(in) :: f(nx, ny, nz)
    real(c_double), intent(out) :: df(nx, ny, nz)
    real(c_double), value :: h
    select case (which)
    case (0);  call derx_00(df, f, h)
    case (1);  call derxp_11(df, f, h)
    case (2);  call derxi_11(df, f, h)
    case (3);  call derxx_00(df, f, h)
    case (4);  call derxxp_11(df, f, h)
    case (5);  call derxxi_11(df, f, h)
    case (6);  call dery_00(df, f, h)
    case (7);  call deryp_11(df, f, h)
    case (8);  call deryi_11(df, f, h)
    case (9);  call deryy_00(df, f, h)
    case (10); call deryyp_11(df, f, h)
    case (11); call deryyi_11(df, f, h)
    case (12); call derz_00(df, f, h)
    case (13); call derzp_11(df, f, h)
    case (14); call derzi_11(df, f, h)
    case (15); call derzz_00(df, f, h)
    case (16); call derzzp_11(df, f, h)
    case (17); call derzzi_11(df, f, h)
    case (18); call derz_2dsim(df, f, h)
    case (19); call derzz_2dsim(df, f, h)
    end select
  end subroutine ref_der

  subroutine ref_divergence(divf, fx, fy, fz, ddx, ddy, ddz, n1, n2, n3, odd) &
       bind(C, name="ref_divergence")
    integer(c_int), value :: n1, n2, n3, odd
    real(c_double), value :: ddx, ddy, ddz
    real(c_double), intent(in) :: fx(n1,n2,n3), fy(n1,n2,n3), fz(n1,n2,n3)
    real(c_double), intent(out) :: divf(n1,n2,n3)
    call divergence(divf, fx, fy, fz, ddx, ddy, ddz, n1, n2, n3, odd)
  end subroutine ref_divergence

  subroutine ref_transeq(ph, u, v, w, sr, fph, rey, schm, a3, b3, c3, it, its, ddx, ddy, ddz, &
       n1, n2, n3, les, nut) bind(C, name="ref_transeq")
    integer(c_int), value :: it, its, n1, n2, n3, les
    real(c_double), value :: rey, schm, ddx, ddy, ddz
    real(c_double), intent(in) :: a3(3), b3(3), c3(3)
    real(c_double), intent(in) :: u(n1,n2,n3), v(n1,n2,n3), w(n1,n2,n3), sr(n1,n2,n3)
    real(c_double), intent(in) :: nut(n1,n2,n3)
    real(c_double), intent(inout) :: ph(n1,n2,n3), fph(n1,n2,n3,3)
    call transeq(ph, u, v, w, sr, fph, rey, schm, a3, b3, c3, it, its, ddx, ddy, ddz, &
         n1, n2, n3, les, nut)
  end subroutine ref_transeq

  subroutine ref_rotational(rx, ry, rz, fx, fy, fz, ddx, ddy, ddz, n1, n2, n3) &
       bind(C, name="ref_rotational")
    integer(c_int), value :: n1, n2, n3
    real(c_double), value :: ddx, ddy, ddz
    real(c_double), intent(in) :: fx(n1,n2,n3), fy(n1,n2,n3), fz(n1,n2,n3)
    real(c_double), intent(out) :: rx(n1,n2,n3), ry(n1,n2,n3), rz(n1,n2,n3)
    call rotational(rx, ry, rz, fx, fy, fz, ddx, ddy, ddz, n1, n2, n3)
  end subroutine ref_rotational

  subroutine ref_q_criterion(q, fx, fy, fz, ddx, ddy, ddz, n1, n2, n3) &
       bind(C, name="ref_q_criterion")
    integer(c_int), value :: n1, n2, n3
    real(c_double), value :: ddx, ddy, ddz
    real(c_double), intent(in) :: fx(n1,n2,n3), fy(n1,n2,n3), fz(n1,n2,n3)
    real(c_double), intent(out) :: q(n1,n2,n3)
    call calculate_Q_criterion(q, fx, fy, fz, ddx, ddy, ddz, n1, n2, n3)
  end subroutine ref_q_criterion

  subroutine ref_nu_t(nut, u, v, w, ddx, ddy, ddz, c_s, dlt, n1, n2, n3) &
       bind(C, name="ref_nu_t")
    integer(c_int), value :: n1, n2, n3
    real(c_double), value :: ddx, ddy, ddz, c_s, dlt
    real(c_double), intent(in) :: u(n1,n2,n3), v(n1,n2,n3), w(n1,n2,n3)
    real(c_double), intent(out) :: nut(n1,n2,n3)
    call calculate_nu_t(nut, u, v, w, ddx, ddy, ddz, c_s, dlt)
  end subroutine ref_nu_t

  subroutine ref_predict(upx, upy, upz, u, v, w, fx, fy, fz, rey, a3, b3, c3, &
       it, its, ddx, ddy, ddz, n1, n2, n3, les, c_s, dlt, nut) &
       bind(C, name="ref_predict")
    integer(c_int), value :: it, its, n1, n2, n3, les
    real(c_double), value :: rey, ddx, ddy, ddz, c_s, dlt
    real(c_double), intent(in) :: a3(3), b3(3), c3(3)
    real(c_double), intent(inout) :: u(n1,n2,n3), v(n1,n2,n3), w(n1,n2,n3)
    real(c_double), intent(inout) :: fx(n1,n2,n3,3), fy(n1,n2,n3,3), fz(n1,n2,n3,3)
    real(c_double), intent(inout) :: nut(n1,n2,n3)
    real(c_double), intent(out) :: upx(n1,n2,n3), upy(n1,n2,n3), upz(n1,n2,n3)
    call predict_velocity(upx, upy, upz, u, v, w, fx, fy, fz, rey, a3, b3, c3, &
         it, its, ddx, ddy, ddz, n1, n2, n3, les, c_s, dlt, nut)
  end subroutine ref_predict

  subroutine ref_correct_pression(p, upx, upy, upz, ddx, ddy, ddz, n1, n2, n3, &
       ddt, om, ep, km, idy) bind(C, name="ref_correct_pression")
    integer(c_int), value :: n1, n2, n3, km, idy
    real(c_double), value :: ddx, ddy, ddz, ddt, ep
    real(c_double), intent(inout) :: om
    real(c_double), intent(inout) :: p(n1,n2,n3)
    real(c_double), intent(in) :: upx(n1,n2,n3), upy(n1,n2,n3), upz(n1,n2,n3)
    call correct_pression(p, upx, upy, upz, ddx, ddy, ddz, n1, n2, n3, &
         ddt, om, ep, km, idy, 0)
  end subroutine ref_correct_pression

  ! Calls the solver bound by schemes() (poisson_solver pointer).
  subroutine ref_poisson(p, rhs, ddx, ddy, ddz, n1, n2, n3, om, ep, km, idy) &
       bind(C, name="ref_poisson")
    integer(c_int), value :: n1, n2, n3, km, idy
    real(c_double), value :: ddx, ddy, ddz, ep
    real(c_double), intent(inout) :: om
    real(c_double), intent(inout) :: p(n1,n2,n3)
    real(c_double), intent(in) :: rhs(n1,n2,n3)
    call poisson_solver(p, rhs, ddx, ddy, ddz, n1, n2, n3, om, ep, km, idy)
  end subroutine ref_poisson

  subroutine ref_correct_velocity(u, v, w, upx, upy, upz, p, ddt, ddx, ddy, ddz, &
       n1, n2, n3) bind(C, name="ref_correct_velocity")
    integer(c_int), value :: n1, n2, n3
    real(c_double), value :: ddt, ddx, ddy, ddz
    real(c_double), intent(inout) :: u(n1,n2,n3), v(n1,n2,n3), w(n1,n2,n3)
    real(c_double), intent(in) :: upx(n1,n2,n3), upy(n1,n2,n3), upz(n1,n2,n3)
    real(c_double), intent(in) :: p(n1,n2,n3)
    call correct_velocity(u, v, w, upx, upy, upz, p, ddt, ddx, ddy, ddz, n1, n2, n3)
  end subroutine ref_correct_velocity

  subroutine ref_function_stats(f, n1, n2, n3, st) bind(C, name="ref_function_stats")
    integer(c_int), value :: n1, n2, n3
    real(c_double), intent(in) :: f(n1,n2,n3)
    real(c_double), intent(out) :: st(6)
    st = function_stats(f, n1, n2, n3)
  end subroutine ref_function_stats

  subroutine ref_statistics(u, v, w, n1, n2, n3, ddx, ddy, ddz, rey, t) &
       bind(C, name="ref_statistics")
    integer(c_int), value :: n1, n2, n3
    real(c_double), value :: ddx, ddy, ddz, rey, t
    real(c_double), intent(in) :: u(n1,n2,n3), v(n1,n2,n3), w(n1,n2,n3)
    call statistics_calc(u, v, w, n1, n2, n3, ddx, ddy, ddz, rey, t)
  end subroutine ref_statistics

  ! AB coefficients exactly as variables() forms them (src/initialization.f90:194-202).
  subroutine ref_ab_coefficients(ddt, a3, b3, c3) bind(C, name="ref_ab_coefficients")
    real(c_double), value :: ddt
    real(c_double), intent(out) :: a3(3), b3(3), c3(3)
    a3(1) = ddt
    b3(1) = 0.d0
    c3(1) = 0.d0
    a3(2) =  3.d0 * ddt / 2.d0
    b3(2) = -1.d0 * ddt / 2.d0
    c3(2) =  0.d0
    a3(3) =  23.d0 * ddt / 12.d0
    b3(3) = -16.d0 * ddt / 12.d0
    c3(3) =   5.d0 * ddt / 12.d0
  end subroutine ref_ab_coefficients

end module o3d_ref_harness
