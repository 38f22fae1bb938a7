! bastr_demo.f90 -- the reference main loop on the GPU from Fortran, through
! the bastr_hip module: the shape mainloop2D/3D (src/solution.F90:36-170)
! takes when its body is replaced by bastr_step.
!
! usage: bastr_demo ia ja ka nsteps seed    (ka = 0 selects 2-D)
! Band forcing at forcek = 5 with target energy 0.075, lproject on, RK3,
! Re = 5000, dt = 1e-4, kmax = 128; the initial field is the seeded hitgen
! field (k0 = 4, <sum u^2> = 1).  Prints the specall and factor_output records
! in the reference's (1X,I20,n(1X,E20.13E2)) format, then the field's mean
! energy.
program bastr_demo
  use, intrinsic :: iso_c_binding
  use bastr_hip
  implicit none
  type(bastr_params) :: p
  type(c_ptr) :: h
  type(bastr_diag), allocatable :: diags(:)
  real(c_double), allocatable :: spec(:), u1(:), u2(:), u3(:), tr(:)
  integer(c_int) :: rc
  integer :: ia, ja, ka, n, s, allkmax
  integer(c_int64_t) :: seed, nstep
  real(c_double) :: time
  character(len=32) :: arg

  call get_command_argument(1, arg); read(arg, *) ia
  call get_command_argument(2, arg); read(arg, *) ja
  call get_command_argument(3, arg); read(arg, *) ka
  call get_command_argument(4, arg); read(arg, *) n
  call get_command_argument(5, arg); read(arg, *) seed

  ! an unsupported grid is refused with a message, as the reference stops
  p = bastr_params(100, 64, 0, 5000.d0, 1.d-4, 1, 2, 5, 0.075d0, 1, 128, BASTR_PREC_F64)
  rc = bastr_create(p, h)
  if (rc /= BASTR_EINVAL) error stop 'expected BASTR_EINVAL for ia = 100'
  write(*, '(A,A)') ' refused: ', bastr_error()

  p = bastr_params(ia, ja, ka, 5000.d0, 1.d-4, 1, 2, 5, 0.075d0, 1, 128, BASTR_PREC_F64)
  if (bastr_create(p, h) /= BASTR_OK) error stop bastr_error()
  if (bastr_init_random_phase(h, 4.d0, seed, 1.d0) /= BASTR_OK) error stop bastr_error()
  allkmax = bastr_allkmax(h)
  allocate(diags(n), spec(3 * allkmax * n))
  rc = bastr_step(h, int(n, c_int32_t), diags, spec)
  if (rc /= BASTR_OK .and. rc /= BASTR_STOP_NAN) error stop bastr_error()
  do s = 1, n
    write(*, '(A,1X,I20,6(1X,E20.13E2))') ' specall', diags(s)%nstep, diags(s)%time, diags(s)%Esspe, &
        diags(s)%Edspe, diags(s)%k2Edspe, diags(s)%kMEdspe
    write(*, '(A,1X,I20,2(1X,E20.13E2))') ' factor ', diags(s)%nstep + 1, diags(s)%force_factor, diags(s)%force_E
  end do
  if (bastr_get_time(h, nstep, time) /= BASTR_OK) error stop bastr_error()
  if (ka == 0) then
    allocate(u1(ia * ja), u2(ia * ja))
    if (bastr_get_fields(h, u1, u2) /= BASTR_OK) error stop bastr_error()
    write(*, '(A,1X,I20,2(1X,E20.13E2))') ' final  ', nstep, time, sum(u1**2 + u2**2) / (ia * ja)
  else
    allocate(u1(ia * ja * ka), u2(ia * ja * ka), u3(ia * ja * ka))
    if (bastr_get_fields(h, u1, u2, u3) /= BASTR_OK) error stop bastr_error()
    write(*, '(A,1X,I20,2(1X,E20.13E2))') ' final  ', nstep, time, sum(u1**2 + u2**2 + u3**2) / (ia * ja * ka)
  end if
  ! transfer2d/3d of the final field, one line per shell as in pp/TransferNNNN.dat
  allocate(tr(3 * allkmax))
  if (ka == 0) then
    rc = bastr_pp_transfer(h, u1, u2, out=tr)
  else
    rc = bastr_pp_transfer(h, u1, u2, u3, out=tr)
  end if
  if (rc /= BASTR_OK) error stop bastr_error()
  do s = 1, allkmax
    write(*, '(A,1X,I20,4(1X,E20.13E2))') ' transf ', nstep, real(s, c_double), tr(s), tr(allkmax + s), &
        tr(2 * allkmax + s)
  end do
  if (bastr_destroy(h) /= BASTR_OK) error stop bastr_error()
end program bastr_demo
