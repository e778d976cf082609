!-----------------------------------------------------------------------
! iso_c_binding interface to libsolar_amd.so (include/solar_amd.h).
!
! The reference calls `RK4` (cube.f90:74, timestep.f90:41) on the COMMON
! /RKWORK/ q(mx,my,mr,5) of one rank; here one call advances all six panels
! held on the GPU, q(nx,ny,nr,5,6) with the panel slowest.
!-----------------------------------------------------------------------
MODULE SOLAR_CAPI

 use iso_c_binding
 implicit none

 integer(c_int), parameter :: SOLAR_OK = 0
 integer(c_int), parameter :: SOLAR_RHS_BC = 1, SOLAR_RHS_SPONGE = 2

 interface

  integer(c_int) function solar_init(ctx, n, nr, gr, c2, rho0, g0, sponge_len, a_sponge) &
                 bind(C, name='solar_init')
   import :: c_int, c_ptr, c_double
   type(c_ptr), intent(out) :: ctx
   integer(c_int), value :: n, nr, sponge_len
   real(c_double), intent(in) :: gr(*), c2(*), rho0(*), g0(*)
   real(c_double), value :: a_sponge
  end function

  integer(c_int) function solar_finalize(ctx) bind(C, name='solar_finalize')
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function

  integer(c_int) function solar_set_state(ctx, q) bind(C, name='solar_set_state')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(in) :: q(*)
  end function

  integer(c_int) function solar_get_state(ctx, q) bind(C, name='solar_get_state')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(out) :: q(*)
  end function

  integer(c_int) function solar_rk4(ctx, nsteps, dt, time) bind(C, name='solar_rk4')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   integer(c_int), value :: nsteps
   real(c_double), value :: dt
   real(c_double), intent(inout) :: time
  end function

  integer(c_int) function solar_compute_rhs(ctx, q, q_rhs, flags) bind(C, name='solar_compute_rhs')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(in) :: q(*)
   real(c_double), intent(out) :: q_rhs(*)
   integer(c_int), value :: flags
  end function

  integer(c_int) function solar_filter_exp(ctx, q, qf) bind(C, name='solar_filter_exp')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(in) :: q(*)
   real(c_double), intent(out) :: qf(*)
  end function

  integer(c_int) function solar_norm_vr(ctx, rms) bind(C, name='solar_norm_vr')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(out) :: rms
  end function

  integer(c_int) function solar_check_finite(ctx) bind(C, name='solar_check_finite')
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function

  integer(c_int) function solar_background(kind, nr, gr, c2, rho0, g0) bind(C, name='solar_background')
   import :: c_int, c_double
   integer(c_int), value :: kind, nr
   real(c_double), intent(out) :: gr(*), c2(*), rho0(*), g0(*)
  end function

  integer(c_int) function solar_generate_ic(kind, n, nr, gr, c2, seed, params, q, info) &
                 bind(C, name='solar_generate_ic')
   import :: c_int, c_double, c_int64_t
   integer(c_int), value :: kind, n, nr
   real(c_double), intent(in) :: gr(*), c2(*), params(*)
   integer(c_int64_t), value :: seed
   real(c_double), intent(out) :: q(*), info(*)
  end function

  ! cube.f90:77 diagnostic: reference_formula /= 0 gives auxiliary.f90:100-125 as written
  integer(c_int) function solar_norm(ctx, field, reference_formula, out) bind(C, name='solar_norm')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   integer(c_int), value :: field, reference_formula
   real(c_double), intent(out) :: out
  end function

  ! WRITE_VISUAL / WRITE_RADIAL_TRACE / WRITE_RESTART (output.f90) of the device state;
  ! dir is a NUL-terminated path (c_null_char alone: the working directory)
  integer(c_int) function solar_write_visual(ctx, dir, cue, o_rad) bind(C, name='solar_write_visual')
   import :: c_int, c_ptr, c_char
   type(c_ptr), value :: ctx
   character(kind=c_char), intent(in) :: dir(*)
   integer(c_int), value :: cue, o_rad
  end function

  integer(c_int) function solar_write_radial_trace(ctx, dir) bind(C, name='solar_write_radial_trace')
   import :: c_int, c_ptr, c_char
   type(c_ptr), value :: ctx
   character(kind=c_char), intent(in) :: dir(*)
  end function

  integer(c_int) function solar_write_restart(ctx, dir, cue, time, step) bind(C, name='solar_write_restart')
   import :: c_int, c_ptr, c_char, c_double
   type(c_ptr), value :: ctx
   character(kind=c_char), intent(in) :: dir(*)
   integer(c_int), value :: cue, step
   real(c_double), value :: time
  end function

  integer(c_int) function solar_read_restart(ctx, dir, cue, time, step) bind(C, name='solar_read_restart')
   import :: c_int, c_ptr, c_char, c_double
   type(c_ptr), value :: ctx
   character(kind=c_char), intent(in) :: dir(*)
   integer(c_int), value :: cue
   real(c_double), intent(out) :: time
   integer(c_int), intent(out) :: step
  end function

  integer(c_int) function solar_output_radius(nr, gr, obsheight, o_rad) bind(C, name='solar_output_radius')
   import :: c_int, c_double
   integer(c_int), value :: nr
   real(c_double), intent(in) :: gr(*)
   real(c_double), value :: obsheight
   integer(c_int), intent(out) :: o_rad
  end function

  ! Sensitivities: 3-D c^2 field, receivers, traces, misfit gradient
  integer(c_int) function solar_set_c2_field(ctx, c2) bind(C, name='solar_set_c2_field')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(in) :: c2(*)
  end function

  integer(c_int) function solar_set_receivers(ctx, nrec, pkji) bind(C, name='solar_set_receivers')
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
   integer(c_int), value :: nrec
   integer(c_int), intent(in) :: pkji(*)
  end function

  integer(c_int) function solar_record(ctx, nsteps, dt, seis) bind(C, name='solar_record')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   integer(c_int), value :: nsteps
   real(c_double), value :: dt
   real(c_double), intent(out) :: seis(*)
  end function

  integer(c_int) function solar_misfit_gradient(ctx, q0, nsteps, dt, data, checkpoint_every, J, grad, seis, &
                                                segment_used) bind(C, name='solar_misfit_gradient')
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(in) :: q0(*), data(*)
   integer(c_int), value :: nsteps, checkpoint_every
   real(c_double), value :: dt
   real(c_double), intent(out) :: J, grad(*), seis(*)
   integer(c_int), intent(out) :: segment_used
  end function

  type(c_ptr) function solar_last_error() bind(C, name='solar_last_error')
   import :: c_ptr
  end function

 end interface

CONTAINS

 ! Stops with the library's message on a nonzero status, the reference's
 ! failure convention (init_mpi.f90:31).
 SUBROUTINE SOLAR_CHECK(rc, what)
  integer(c_int), intent(in) :: rc
  character(len=*), intent(in) :: what
  character(kind=c_char), pointer :: msg(:)
  integer :: i
  character(len=512) :: text
  if (rc == SOLAR_OK) return
  call c_f_pointer(solar_last_error(), msg, [512])
  text = ' '
  do i = 1, 512
   if (msg(i) == c_null_char) exit
   text(i:i) = msg(i)
  enddo
  print *, 'solar: ', what, ' failed (', rc, '): ', trim(text)
  stop 1
 END SUBROUTINE SOLAR_CHECK

END MODULE SOLAR_CAPI
