! Drop-in module `derivation` (reference: src/derivation.f90).
!
! Same module name, procedure names and dummy arguments as the reference, so
! the reference's schemes() (src/initialization.f90:226-304) binds its
! procedure pointers (interface der_type, :86-91) to these routines unchanged.
! Each body runs the corresponding HIP stencil kernel of libo3d.so through
! o3d_der (include/o3d.h); results are bit-identical to the reference's.
module derivation
  use o3d_capi
  implicit none
  private :: run_der, run_der_x

contains

  subroutine run_der(which, df, f, d)
    integer, intent(in) :: which
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: d
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der_x(which, size(f, 1), size(f, 2), size(f, 3), df, f, d)
  end subroutine run_der

  ! Explicit-shape dummies: a non-contiguous actual argument is copied in and
  ! out by the compiler, a contiguous one is passed in place.
  subroutine run_der_x(which, n1, n2, n3, df, f, d)
    integer, intent(in) :: which, n1, n2, n3
    real(kind=8), intent(in) :: f(n1, n2, n3)
    real(kind=8), intent(in) :: d
    real(kind=8), intent(out) :: df(n1, n2, n3)
    call o3d_check(o3d_der(int(which, c_int), df, f, int(n1, c_int), int(n2, c_int), &
         int(n3, c_int), d), 'derivation')
  end subroutine run_der_x

  !> X first derivative, periodic (src/derivation.f90:6-60)
  subroutine derx_00(df, f, dx)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dx
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(0, df, f, dx)
  end subroutine derx_00

  !> X first derivative, free-slip even (src/derivation.f90:62-109)
  subroutine derxp_11(df, f, dx)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dx
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(1, df, f, dx)
  end subroutine derxp_11

  !> X first derivative, free-slip odd (src/derivation.f90:111-163)
  subroutine derxi_11(df, f, dx)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dx
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(2, df, f, dx)
  end subroutine derxi_11

  !> Y first derivative, periodic (src/derivation.f90:165-217)
  subroutine dery_00(df, f, dy)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dy
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(6, df, f, dy)
  end subroutine dery_00

  !> Y first derivative, free-slip even (src/derivation.f90:219-267)
  subroutine deryp_11(df, f, dy)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dy
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(7, df, f, dy)
  end subroutine deryp_11

  !> Y first derivative, free-slip odd (src/derivation.f90:269-321)
  subroutine deryi_11(df, f, dy)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dy
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(8, df, f, dy)
  end subroutine deryi_11

  !> Z first derivative, periodic (src/derivation.f90:323-375)
  subroutine derz_00(df, f, dz)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dz
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(12, df, f, dz)
  end subroutine derz_00

  !> Z first derivative, free-slip even (src/derivation.f90:377-425)
  subroutine derzp_11(df, f, dz)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dz
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(13, df, f, dz)
  end subroutine derzp_11

  !> Z first derivative, free-slip odd (src/derivation.f90:427-479)
  subroutine derzi_11(df, f, dz)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dz
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(14, df, f, dz)
  end subroutine derzi_11

  !> Z first derivative of a 2D run: 0*f*dz (src/derivation.f90:481-495)
  subroutine derz_2dsim(df, f, dz)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dz
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(18, df, f, dz)
  end subroutine derz_2dsim

  !> X second derivative, periodic (src/derivation.f90:497-542)
  subroutine derxx_00(df, f, dx)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dx
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(3, df, f, dx)
  end subroutine derxx_00

  !> X second derivative, free-slip odd (src/derivation.f90:544-591)
  subroutine derxxi_11(df, f, dx)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dx
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(5, df, f, dx)
  end subroutine derxxi_11

  !> X second derivative, free-slip even (src/derivation.f90:593-640)
  subroutine derxxp_11(df, f, dx)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dx
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(4, df, f, dx)
  end subroutine derxxp_11

  !> Y second derivative, periodic (src/derivation.f90:642-689)
  subroutine deryy_00(df, f, dy)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dy
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(9, df, f, dy)
  end subroutine deryy_00

  !> Y second derivative, free-slip even (src/derivation.f90:691-738)
  subroutine deryyp_11(df, f, dy)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dy
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(10, df, f, dy)
  end subroutine deryyp_11

  !> Y second derivative, free-slip odd (src/derivation.f90:740-787)
  subroutine deryyi_11(df, f, dy)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dy
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(11, df, f, dy)
  end subroutine deryyi_11

  !> Z second derivative, periodic (src/derivation.f90:789-834)
  subroutine derzz_00(df, f, dz)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dz
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(15, df, f, dz)
  end subroutine derzz_00

  !> Z second derivative, free-slip odd (src/derivation.f90:836-883)
  subroutine derzzi_11(df, f, dz)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dz
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(17, df, f, dz)
  end subroutine derzzi_11

  !> Z second derivative, free-slip even (src/derivation.f90:885-932)
  subroutine derzzp_11(df, f, dz)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dz
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(16, df, f, dz)
  end subroutine derzzp_11

  !> Z second derivative of a 2D run: 0*f*dz (src/derivation.f90:934-948)
  subroutine derzz_2dsim(df, f, dz)
    real(kind=8), intent(in) :: f(:,:,:)
    real(kind=8), intent(in) :: dz
    real(kind=8), intent(out) :: df(:,:,:)
    call run_der(19, df, f, dz)
  end subroutine derzz_2dsim

  !> 1D first derivative used by the initial conditions (src/derivation.f90:950-992):
  !> the 6th-order central stencil inside, one-sided 2nd-order differences at the
  !> two end nodes and central 2nd/4th-order differences next to them.  Runs on
  !> the host: it is called once per run on a single line of nodes.
  subroutine dery1D(df, f, dy)
    real(kind=8), intent(in) :: f(:)
    real(kind=8), intent(in) :: dy
    real(kind=8), intent(out) :: df(:)
    real(kind=8) :: h60, c1, c2, c3
    integer :: n, j
    n = size(f)
    h60 = 60.d0 * dy
    c1 = 1.d0 / h60
    c2 = 9.d0 / h60
    c3 = 45.d0 / h60
    do j = 4, n - 3
       df(j) = c1 * (f(j+3) - f(j-3)) - c2 * (f(j+2) - f(j-2)) + c3 * (f(j+1) - f(j-1))
    end do
    df(1) = (-f(3) + 4.d0 * f(2) - 3.d0 * f(1)) / (2.d0 * dy)
    df(n) = (3.d0 * f(n) - 4.d0 * f(n-1) + f(n-2)) / (2.d0 * dy)
    df(2) = (f(3) - f(1)) / (2.d0 * dy)
    df(n-1) = (f(n) - f(n-2)) / (2.d0 * dy)
    df(3) = (-f(5) + 8.d0 * f(4) - 8.d0 * f(2) + f(1)) / (12.d0 * dy)
    df(n-2) = (-f(n) + 8.d0 * f(n-1) - 8.d0 * f(n-3) + f(n-4)) / (12.d0 * dy)
  end subroutine dery1D

end module derivation
