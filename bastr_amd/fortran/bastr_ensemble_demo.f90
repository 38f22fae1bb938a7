! bastr_ensemble_demo: a Fortran host driving an ensemble of 2-D runs through
! the bastr_hip module (bastr_create_ensemble, the member entry points and
! bastr_step_group).  Usage: bastr_ensemble_demo N NSTEPS
!   a 2-member N x N ensemble (member 0 band-forced, member 1 free decay, the
!   seeded generator with k0 = 4, seeds 1 and 2) and a one-field N x N run
!   (seed 3), stepped together; prints one "ens m i Esspe Edspe" line per
!   member and iteration, "one i Esspe Edspe" for the one-field run, and
!   "time m nstep time" per member at the end.
program bastr_ensemble_demo
  use iso_c_binding
  use bastr_hip
  implicit none
  type(bastr_params) :: p
  type(c_ptr), target :: hs(2)
  type(c_ptr) :: dp(2)
  type(bastr_diag), allocatable, target :: de(:), d1(:)
  integer(c_int32_t) :: forced(2), n, nsteps, m, i
  integer(c_int64_t) :: ns
  real(c_double) :: t
  integer :: rc
  character(len=32) :: arg

  call get_command_argument(1, arg); read (arg, *) n
  call get_command_argument(2, arg); read (arg, *) nsteps
  p%ia = n; p%ja = n; p%ka = 0
  p%reynolds = 5000.0d0; p%deltat = 1.0d-4
  p%timemethod = 1; p%forcemethod = 2; p%forcek = 5; p%target_energy = 0.075d0
  p%lproject = 1; p%kmax = 128; p%precision = 0
  forced = [1, 0]
  call check(bastr_create_ensemble(p, 2_c_int32_t, forced, hs(1)))
  if (bastr_nmembers(hs(1)) /= 2) error stop 'nmembers'
  do m = 0, 1
    call check(bastr_init_random_phase_member(hs(1), m, 4.0d0, int(m + 1, c_int64_t), 1.0d0))
  end do
  call check(bastr_create(p, hs(2)))
  call check(bastr_init_random_phase(hs(2), 4.0d0, 3_c_int64_t, 1.0d0))
  allocate (de(2 * nsteps), d1(nsteps))
  dp(1) = c_loc(de)
  dp(2) = c_loc(d1)
  call check(bastr_step_group(hs, 2_c_int32_t, nsteps, diags=dp))
  do m = 0, 1
    do i = 1, nsteps
      write (*, '(A,1X,I0,1X,I0,2(1X,ES24.16E3))') 'ens', m, i - 1, de(m * nsteps + i)%Esspe, &
        de(m * nsteps + i)%Edspe
    end do
  end do
  do i = 1, nsteps
    write (*, '(A,1X,I0,2(1X,ES24.16E3))') 'one', i - 1, d1(i)%Esspe, d1(i)%Edspe
  end do
  do m = 0, 1
    call check(bastr_get_member_time(hs(1), m, ns, t))
    write (*, '(A,1X,I0,1X,I0,1X,ES24.16E3)') 'time', m, ns, t
  end do
  rc = bastr_destroy(hs(1))
  rc = bastr_destroy(hs(2))

contains

  subroutine check(code)
    integer(c_int), intent(in) :: code
    if (code /= 0) then
      write (*, '(A)') 'error: '//bastr_error()
      error stop 1
    end if
  end subroutine

end program
