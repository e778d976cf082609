!-----------------------------------------------------------------------
! Fortran driver of the GPU hot path.  Same call sequence as the reference
! driver (cube.f90:35-92): background (SOLAR_DATA), set-up of metric and
! interpolation tables, time step, initial condition, then the time loop of
! RK4 calls with the per-step print of step, dt, time, the v_r norm and the
! step wall time, and every 6th step the visual and radial-trace output.
! The numerics below RK4 run on the GPU through SOLAR_CAPI.
!
! usage: cube_gpu n nr background ic nsteps seed [state_out [outdir]]
!   background 0 uniform / 1 stratified, ic 0 Gaussian pulse / 1 Y_lm mode
!   state_out: file that receives the final q(n,n,nr,5,6) as raw float64
!   outdir: every 6th step WRITE_VISUAL and WRITE_RADIAL_TRACE there, as
!           cube.f90:84-89 (obsheight 0.98); '-' or absent: no output
!-----------------------------------------------------------------------
PROGRAM CUBE_GPU

 use iso_c_binding
 use SOLAR_CAPI
 implicit none

 integer(c_int) :: n, nr, bg_kind, ic_kind, nsteps, sponge_len, rc, o_rad, cue_vis
 integer(c_int64_t) :: seed
 real(c_double), allocatable :: gr(:), c2(:), rho0(:), g0(:), q(:)
 real(c_double) :: params(3), info(4), dt, time, dxi, rms, tt, a_sponge
 type(c_ptr) :: ctx
 integer :: step, nargs, u
 integer(8) :: t0, t1, rate
 character(len=256) :: arg, state_out, outdir
 character(kind=c_char, len=257) :: cdir

 nargs = command_argument_count()
 if (nargs < 6) then
  print *, 'usage: cube_gpu n nr background ic nsteps seed [state_out]'
  stop 2
 endif
 call get_command_argument(1, arg); read(arg, *) n
 call get_command_argument(2, arg); read(arg, *) nr
 call get_command_argument(3, arg); read(arg, *) bg_kind
 call get_command_argument(4, arg); read(arg, *) ic_kind
 call get_command_argument(5, arg); read(arg, *) nsteps
 call get_command_argument(6, arg); read(arg, *) seed
 state_out = ' '
 if (nargs >= 7) call get_command_argument(7, state_out)
 if (trim(state_out) == '-') state_out = ' '
 outdir = ' '
 if (nargs >= 8) call get_command_argument(8, outdir)
 if (trim(outdir) == '-') outdir = ' '
 cdir = trim(outdir) // c_null_char

 ! SOLAR_DATA (cube.f90:37): nondimensional radial background
 allocate(gr(nr), c2(nr), rho0(nr), g0(nr))
 rc = solar_background(bg_kind, nr, gr, c2, rho0, g0)
 call SOLAR_CHECK(rc, 'solar_background')

 ! SETUP_COORDINATES_METRICS_ALL, SETUP_LAGRANGE_INTERPOLANT (cube.f90:39-43)
 sponge_len = max(1, nint(0.08d0 * nr))
 a_sponge = 0.1d0
 rc = solar_init(ctx, n, nr, gr, c2, rho0, g0, sponge_len, a_sponge)
 call SOLAR_CHECK(rc, 'solar_init')

 ! COMPUTE_TIMESTEP (cube.f90:45): fixed CFL-0.5 step, as the workloads define it
 dxi = (acos(-1d0) / 2d0) / (n - 1)
 dt = 0.5d0 * min(gr(2) - gr(1), gr(1) * dxi) / sqrt(maxval(c2))

 ! GENERATE_IC (cube.f90:47)
 allocate(q(int(n, 8) * n * nr * 30))
 if (ic_kind == 0) then
  params = [1d-3, 0.05d0, 0d0]
 else
  params = [1d-3, 1d-6, 20d0]
 endif
 rc = solar_generate_ic(ic_kind, n, nr, gr, c2, seed, params, q, info)
 call SOLAR_CHECK(rc, 'solar_generate_ic')
 rc = solar_set_state(ctx, q)
 call SOLAR_CHECK(rc, 'solar_set_state')
 print *, 'nstep = ', nsteps
 ! output height (ic_acoustic_wave.f90:128-133)
 rc = solar_output_radius(nr, gr, 0.98d0, o_rad)
 call SOLAR_CHECK(rc, 'solar_output_radius')
 cue_vis = 0

 time = 0d0
 call system_clock(count_rate=rate)
 do step = 0, nsteps - 1
  call system_clock(t0)
  rc = solar_rk4(ctx, 1, dt, time)              ! call RK4 (cube.f90:74)
  call SOLAR_CHECK(rc, 'solar_rk4')
  call system_clock(t1)
  rc = solar_norm_vr(ctx, rms)
  call SOLAR_CHECK(rc, 'solar_norm_vr')
  rc = solar_norm(ctx, 3, 1, tt)                 ! tt = norm(q(:,:,:,4)), cube.f90:77
  call SOLAR_CHECK(rc, 'solar_norm')
  if (mod(step, max(1, nsteps / 10)) == 0 .or. step == nsteps - 1) then
   print *, step, dt, time
   print *, tt, rms
   print *, 'steptime = ', dble(t1 - t0) / dble(rate)
  endif
  if (len_trim(outdir) > 0 .and. mod(step, 6) == 0) then   ! cube.f90:84-89
   rc = solar_write_visual(ctx, cdir, cue_vis, o_rad)
   call SOLAR_CHECK(rc, 'solar_write_visual')
   rc = solar_write_radial_trace(ctx, cdir)
   call SOLAR_CHECK(rc, 'solar_write_radial_trace')
   cue_vis = cue_vis + 1
  endif
 enddo

 rc = solar_check_finite(ctx)
 call SOLAR_CHECK(rc, 'solar_check_finite')
 if (len_trim(state_out) > 0) then
  rc = solar_get_state(ctx, q)
  call SOLAR_CHECK(rc, 'solar_get_state')
  open(newunit=u, file=trim(state_out), access='stream', form='unformatted', status='replace')
  write(u) q
  close(u)
 endif
 rc = solar_finalize(ctx)
 deallocate(gr, c2, rho0, g0, q)

END PROGRAM CUBE_GPU
