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
