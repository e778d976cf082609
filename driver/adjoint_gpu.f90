!-----------------------------------------------------------------------
! Fortran driver of the sensitivity path: receiver misfit of the c^2 field
! and its gradient by the discrete adjoint, through SOLAR_CAPI.
!
! usage: adjoint_gpu n nr nsteps [grad_out]
!   Stratified background, spherical-harmonic initial state (as the c5
!   workload); 8 receivers of p' near mid-shell on all six panels; data from
!   a model whose c^2 carries a 2% radial bump.  Prints J, |grad|, and the directional derivative along
!   the bump by the adjoint and by a central difference of J.
!   grad_out: file that receives dJ/dc2(n,n,nr,6) as raw float64
!-----------------------------------------------------------------------
PROGRAM ADJOINT_GPU

 use iso_c_binding
 use SOLAR_CAPI
 implicit none

 integer(c_int) :: n, nr, nsteps, sponge_len, rc, nrec, seg
 real(c_double), allocatable :: gr(:), c2(:), rho0(:), g0(:), q0(:), c2f(:), c2t(:), dirn(:)
 real(c_double), allocatable :: data(:), seis(:), grad(:)
 real(c_double) :: params(3), info(4), dt, dxi, J, Jp, Jm, eps, fd, ad, time
 integer(c_int), allocatable :: rec(:)
 type(c_ptr) :: ctx
 integer :: i, k, p, r, nargs, u
 integer(8) :: nf
 character(len=256) :: arg, grad_out

 nargs = command_argument_count()
 if (nargs < 3) then
  print *, 'usage: adjoint_gpu n nr nsteps [grad_out]'
  stop 2
 endif
 call get_command_argument(1, arg); read(arg, *) n
 call get_command_argument(2, arg); read(arg, *) nr
 call get_command_argument(3, arg); read(arg, *) nsteps
 grad_out = ' '
 if (nargs >= 4) call get_command_argument(4, grad_out)

 allocate(gr(nr), c2(nr), rho0(nr), g0(nr))
 rc = solar_background(1, nr, gr, c2, rho0, g0)
 call SOLAR_CHECK(rc, 'solar_background')
 sponge_len = max(1, nint(0.08d0 * nr))
 rc = solar_init(ctx, n, nr, gr, c2, rho0, g0, sponge_len, 0.1d0)
 call SOLAR_CHECK(rc, 'solar_init')
 dxi = (acos(-1d0) / 2d0) / (n - 1)
 dt = 0.5d0 * min(gr(2) - gr(1), gr(1) * dxi) / sqrt(maxval(c2))

 nf = int(n, 8) * n * nr
 allocate(q0(nf * 30), c2f(nf * 6), c2t(nf * 6), dirn(nf * 6), grad(nf * 6))
 params = [1d-3, 1d-6, 20d0]
 rc = solar_generate_ic(1, n, nr, gr, c2, 7_c_int64_t, params, q0, info)
 call SOLAR_CHECK(rc, 'solar_generate_ic')

 ! c2 fields (n,n,nr,6): the radial profile, and the "true" model with a bump
 do p = 0, 5
  do k = 1, nr
   do i = 1, n * n
    c2f(p * nf + (k - 1) * n * n + i) = c2(k)
    dirn(p * nf + (k - 1) * n * n + i) = c2(k) * exp(-((k - nr / 2d0) / (nr / 8d0))**2)
   enddo
  enddo
 enddo
 c2t = c2f + 0.02d0 * dirn

 ! receivers (panel, k, j, i), 0-based
 nrec = 8
 allocate(rec(4 * nrec), data(nsteps * nrec), seis(nsteps * nrec))
 do r = 0, nrec - 1
  rec(4 * r + 1) = mod(r, 6)
  rec(4 * r + 2) = nr / 2 + mod(r, 3) - 1
  rec(4 * r + 3) = n / 2 - 1 + mod(r, 2)
  rec(4 * r + 4) = mod(5 * r + 2, n)
 enddo
 rc = solar_set_receivers(ctx, nrec, rec)
 call SOLAR_CHECK(rc, 'solar_set_receivers')

 ! data from the true model
 rc = solar_set_c2_field(ctx, c2t); call SOLAR_CHECK(rc, 'solar_set_c2_field')
 rc = solar_set_state(ctx, q0); call SOLAR_CHECK(rc, 'solar_set_state')
 rc = solar_record(ctx, nsteps, dt, data); call SOLAR_CHECK(rc, 'solar_record')

 ! misfit and gradient at the radial model
 rc = solar_set_c2_field(ctx, c2f); call SOLAR_CHECK(rc, 'solar_set_c2_field')
 rc = solar_misfit_gradient(ctx, q0, nsteps, dt, data, 0, J, grad, seis, seg)
 call SOLAR_CHECK(rc, 'solar_misfit_gradient')
 ad = sum(grad * dirn)

 ! central difference of J along the bump
 eps = 1d-4
 Jp = misfit(c2f + eps * dirn)
 Jm = misfit(c2f - eps * dirn)
 fd = (Jp - Jm) / (2d0 * eps)
 print *, 'J = ', J, ' |grad| = ', sqrt(sum(grad**2)), ' segment = ', seg
 print *, 'dJ adjoint = ', ad, ' dJ central difference = ', fd, ' rel = ', abs(ad - fd) / abs(fd)
 if (len_trim(grad_out) > 0) then
  open(newunit=u, file=trim(grad_out), access='stream', form='unformatted', status='replace')
  write(u) grad
  close(u)
 endif
 rc = solar_finalize(ctx)
 time = 0d0

CONTAINS

 real(c_double) function misfit(field)
  real(c_double), intent(in) :: field(:)
  integer(c_int) :: rc2
  rc2 = solar_set_c2_field(ctx, field); call SOLAR_CHECK(rc2, 'solar_set_c2_field')
  rc2 = solar_set_state(ctx, q0); call SOLAR_CHECK(rc2, 'solar_set_state')
  rc2 = solar_record(ctx, nsteps, dt, seis); call SOLAR_CHECK(rc2, 'solar_record')
  misfit = sum((seis - data)**2)
 end function

END PROGRAM ADJOINT_GPU
