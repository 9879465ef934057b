! Test infrastructure only (the oracle and the CPU baseline, never the product path).
!
! Runs the reference osinco3d time loop on the CPU with the reference's own
! modules, compiled from the reference sources in place by oracle/Makefile.
! It follows the step sequence of src/osinco3d_main.f90:97-188 (old_values,
! predict_velocity, correct_pression, correct_velocity, the two diagnostic
! divergences with function_stats, print_velocity_values, compute_cfl, and
! every 25 steps calculate_residuals and statistics_calc, which appends to
! outputs/stats.dat when the run directory has outputs/) but leaves out the
! rest of the driver's file output (slices, XDMF, checkpoints, gnuplot),
! which needs no timing and writes outside the run directory.
! The full driver cannot be built here: src/initial_conditions.f90 needs
! src/noise_module.f90, which needs FFTW3, absent from this image.
!
! Usage (run directory holds parameters.o3d, read by the reference's own
! parameters() from src/initialization.f90:114-151):
!   o3d_ref_run <out.bin | -> [dump_every]      ('-': no field output)
! typesim = 2 : Taylor-Green initial field, same formula as
!               src/initial_conditions.f90:157-162 (restated here because that
!               module cannot be built without FFTW);
! typesim = 0 : initial field read by read_fields (src/IOfunctions.f90:406-470)
!               from fields.bin.
! Output: stream file of records [itime(int32), nx, ny, nz (int32), time,
! ux, uy, uz, pp (float64, column-major)]: the initial field (itime 0), then
! every step that is a multiple of dump_every, and the last step.
! Per-step wall time goes to stdout as "O3D_STEP_SECONDS <itime> <seconds>",
! and per phase as "O3D_PHASE_SECONDS <itime> <predict> <correct_pression>
! <correct_velocity> <one divergence> <remaining diagnostics>".  Built as
! the drop-in (O3D_DROPIN: the same loop over this package's modules), it
! also prints the host <-> device bytes of each step as
! "O3D_STEP_TRANSFER <itime> <h2d> <d2h>".
program o3d_ref_run
  use initialization
  use integration
  use diffoper
  use functions
  use utils
  use IOfunctions
#ifdef O3D_DROPIN
  use o3d_capi, only: o3d_transfer_stats
  use iso_c_binding, only: c_double, c_int
#endif
  implicit none
#ifdef O3D_DROPIN
  real(c_double) :: h2d, d2h
  integer(c_int) :: trc
#endif
  character(len=256) :: outname, arg
  integer :: dump_every, i, j, k, iunit
  integer(8) :: c0, c1, crate, cp(6)
  real(kind=8) :: t_step, tph(5)
  logical :: dump

  call get_command_argument(1, outname)
  dump_every = 0
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg)
     read(arg, *) dump_every
  end if

  call parameters()
  call variables()
  call schemes()

  if (typesim == 2) then
     do k = 1, nz
        do j = 1, ny
           do i = 1, nx
              ux(i,j,k) =  ratio * u0/l0 * sin(x(i)) * cos(y(j))* cos(z(k))
              uy(i,j,k) = -ratio * u0/l0 * cos(x(i)) * sin(y(j))* cos(z(k))
              uz(i,j,k) = 0.d0
              pp(i,j,k) = 0.0625d0 * (cos(2.d0*x(i)) + cos(2.d0*y(j))) * &
                   (cos(2.d0*z(k)) + 2.d0)
           end do
        end do
     end do
     time0 = 0.d0
  else
     fields_name_file = 'fields.bin'
     call read_fields(x, y, z, ux, uy, uz, pp, phi, nx, ny, nz, time0, fields_name_file)
     time = time0
     phi = 0.d0
  end if
  nu_t = 0.d0

  dump = trim(outname) /= '-'
  if (dump) then
     open(newunit=iunit, file=trim(outname), status='replace', access='stream', &
          form='unformatted', action='write')
     write(iunit) int(0, 4), int(nx, 4), int(ny, 4), int(nz, 4)
     write(iunit) time0
     write(iunit) ux, uy, uz, pp
  end if
  call system_clock(count_rate=crate)
  do itime = itstart, itstop
     call system_clock(c0)
     cp(1) = c0
     time = time0 + itime * dt
     call old_values(ux, uy, uz, old_ux, old_uy, old_uz, nx, ny, nz)
     call predict_velocity(ux_pred, uy_pred, uz_pred, ux, uy, uz, &
          fux, fuy, fuz, re, adt, bdt, cdt, itime, itscheme, &
          dx, dy, dz, nx, ny, nz, iles, cs, delta, nu_t)
     call system_clock(cp(2))
     call correct_pression(pp, ux_pred, uy_pred, uz_pred, dx, dy, dz, &
          nx, ny, nz, dt, omega, eps, kmax, idyn, multigrid)
     call system_clock(cp(3))
     call correct_velocity(ux, uy, uz, ux_pred, uy_pred, uz_pred, &
          pp, dt, dx, dy, dz, nx, ny, nz)
     call system_clock(cp(4))
     call divergence(divu_pred, ux_pred, uy_pred, uz_pred, &
          dx, dy, dz, nx, ny, nz, 1)
     call system_clock(cp(5))
     divu_stats = function_stats(divu_pred, nx, ny, nz)
     call print_divu_statistics(divu_stats, .true.)
     call divergence(divu, ux, uy, uz, dx, dy, dz, nx, ny, nz, 1)
     divu_stats = function_stats(divu, nx, ny, nz)
     call print_divu_statistics(divu_stats, .false.)
     call print_velocity_values(ux, uy, uz)
     call compute_cfl(cflx, cfly, cflz, ux, uy, uz, dx, dy, dz, dt)
     call print_cfl(cflx, cfly, cflz)
     if (modulo(itime, 25) == 0) then
        ! src/osinco3d_main.f90:167-181 (without the profile files)
        call calculate_residuals(ux, uy, uz, old_ux, old_uy, old_uz, dt, &
             t_ref, u_ref, nx, ny, nz, itime)
        if (time > initstat) then
           call statistics_calc(ux, uy, uz, nx, ny, nz, dx, dy, dz, re, time)
        end if
     end if
     call system_clock(c1)
     cp(6) = c1
     t_step = real(c1 - c0, kind=8) / real(crate, kind=8)
     do i = 1, 5
        tph(i) = real(cp(i+1) - cp(i), kind=8) / real(crate, kind=8)
     end do
     write(*, '(A17,I8,ES16.8)') 'O3D_STEP_SECONDS ', itime, t_step
     ! predict, correct_pression, correct_velocity, one divergence, rest of
     ! the diagnostics
     write(*, '(A18,I8,5ES16.8)') 'O3D_PHASE_SECONDS ', itime, tph
#ifdef O3D_DROPIN
     trc = o3d_transfer_stats(h2d, d2h, 1_c_int)
     write(*, '(A18,I8,2ES16.8)') 'O3D_STEP_TRANSFER ', itime, h2d, d2h
#endif
     if (dump .and. ((dump_every > 0 .and. modulo(itime, max(dump_every, 1)) == 0) .or. &
          itime == itstop)) then
        write(iunit) int(itime, 4), int(nx, 4), int(ny, 4), int(nz, 4)
        write(iunit) time
        write(iunit) ux, uy, uz, pp
        flush(iunit)
     end if
  end do
  if (dump) close(iunit)
end program o3d_ref_run
