! bastr_hip.f90 -- ISO_C_BINDING interface to the MI355X step library
! (include/bastr.h), the module a Fortran Bastr build uses in place of the
! FFT boundary module fftwlink (src/fftwlink.F90) and the solver loop it
! serves (src/solution.F90).
!
! Every function returns 0 (BASTR_OK) on success; a Fortran caller turns a
! nonzero return into `error stop bastr_error()`, the fail-fast contract of
! the reference's `stop '...'` (fftwlink.F90:79,300,319,419).
module bastr_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: bastr_params, bastr_diag, bastr_error
  public :: bastr_create, bastr_destroy, bastr_allkmax, bastr_set_fields, bastr_get_fields
  public :: bastr_init_random_phase, bastr_set_time, bastr_get_time, bastr_step
  public :: bastr_shell_counts, bastr_fft_forward, bastr_fft_backward
  public :: bastr_pp_spectra, bastr_pp_transfer
  public :: bastr_create_ensemble, bastr_nmembers, bastr_set_member_fields, bastr_get_member_fields
  public :: bastr_init_random_phase_member, bastr_get_member_time, bastr_step_group
  public :: bastr_set_spectrum_law, bastr_step_noalloc

  integer(c_int), parameter, public :: BASTR_OK = 0, BASTR_STOP_NAN = 1
  integer(c_int), parameter, public :: BASTR_EINVAL = -1, BASTR_EDEVICE = -2, BASTR_ENOMEM = -3
  integer(c_int32_t), parameter, public :: BASTR_PREC_F64 = 0, BASTR_PREC_F32 = 1
  integer(c_int32_t), parameter, public :: BASTR_LAW_HITGEN = 0, BASTR_LAW_PURE = 1

  ! commvar.F90:16-66 run parameters (bastr.h: bastr_params)
  type, bind(C) :: bastr_params
    integer(c_int32_t) :: ia, ja, ka
    real(c_double)     :: reynolds, deltat
    integer(c_int32_t) :: timemethod, forcemethod, forcek
    real(c_double)     :: target_energy
    integer(c_int32_t) :: lproject, kmax, precision
  end type

  ! one loop iteration's log values (bastr.h: bastr_diag)
  type, bind(C) :: bastr_diag
    integer(c_int64_t) :: nstep
    real(c_double) :: time, Esspe, Edspe, k2Edspe, kMEdspe
    real(c_double) :: umumtheta2, umumijji, u2theta, dissp, eta_min, Taylength, ReTay, Kollength
    real(c_double) :: force_factor, force_E, force_energy, force_Fen
    integer(c_int32_t) :: nan_flag, pad_
  end type

  interface
    type(c_ptr) function bastr_last_error() bind(C, name='bastr_last_error')
      import
    end function
    integer(c_int) function bastr_create(p, h) bind(C, name='bastr_create')
      import
      type(bastr_params), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function bastr_destroy(h) bind(C, name='bastr_destroy')
      import
      type(c_ptr), value :: h
    end function
    integer(c_int) function bastr_allkmax(h) bind(C, name='bastr_allkmax')
      import
      type(c_ptr), value :: h
    end function
    ! u3 is absent in 2-D (a NULL pointer on the C side)
    integer(c_int) function bastr_set_fields(h, u1, u2, u3) bind(C, name='bastr_set_fields')
      import
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u1(*), u2(*)
      real(c_double), intent(in), optional :: u3(*)
    end function
    integer(c_int) function bastr_get_fields(h, u1, u2, u3) bind(C, name='bastr_get_fields')
      import
      type(c_ptr), value :: h
      real(c_double), intent(out) :: u1(*), u2(*)
      real(c_double), intent(out), optional :: u3(*)
    end function
    integer(c_int) function bastr_init_random_phase(h, k0, seed, energy) bind(C, name='bastr_init_random_phase')
      import
      type(c_ptr), value :: h
      real(c_double), value :: k0, energy
      integer(c_int64_t), value :: seed
    end function
    integer(c_int) function bastr_set_spectrum_law(h, law) bind(C, name='bastr_set_spectrum_law')
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: law
    end function
    integer(c_int) function bastr_set_time(h, nstep, time) bind(C, name='bastr_set_time')
      import
      type(c_ptr), value :: h
      integer(c_int64_t), value :: nstep
      real(c_double), value :: time
    end function
    integer(c_int) function bastr_get_time(h, nstep, time) bind(C, name='bastr_get_time')
      import
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: nstep
      real(c_double), intent(out) :: time
    end function
    integer(c_int) function bastr_step(h, n, diags, spectra) bind(C, name='bastr_step')
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      type(bastr_diag), intent(out) :: diags(*)
      real(c_double), intent(out), optional :: spectra(*)
    end function
    integer(c_int) function bastr_shell_counts(h, counts) bind(C, name='bastr_shell_counts')
      import
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: counts(*)
    end function
    ! fft2d/fft3d and ifft2d/ifft3d of fftwlink.F90:324-368 on host arrays
    integer(c_int) function bastr_fft_forward(h, a) bind(C, name='bastr_fft_forward')
      import
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: a(*)
    end function
    integer(c_int) function bastr_fft_backward(h, a) bind(C, name='bastr_fft_backward')
      import
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: a(*)
    end function
    ! Bastrpp spectra2d (2-D): shells = Es(1:allkmax), Ed(1:allkmax);
    ! totals = Esspe, Edspe, k2Edspe, kM1Edspe.  Pass shells/totals by keyword.
    integer(c_int) function bastr_pp_spectra(h, u1, u2, u3, shells, totals) bind(C, name='bastr_pp_spectra')
      import
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u1(*), u2(*)
      real(c_double), intent(in), optional :: u3(*)
      real(c_double), intent(out) :: shells(*), totals(4)
    end function
    ! Bastrpp transfer2d/3d: out = T(1:allkmax), T1(1:allkmax), T2(1:allkmax)
    integer(c_int) function bastr_pp_transfer(h, u1, u2, u3, out) bind(C, name='bastr_pp_transfer')
      import
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u1(*), u2(*)
      real(c_double), intent(in), optional :: u3(*)
      real(c_double), intent(out) :: out(*)
    end function
    ! Ensembles of nmembers independent 2-D runs of one grid (no counterpart in
    ! the reference): forced(m) = 0 free decay, 1 forced as p says (absent: all
    ! forced); members are numbered from 0.  bastr_step on an ensemble fills
    ! diags(m*n + i) and spectra(m*n*3*allkmax + ...) for member m.
    integer(c_int) function bastr_create_ensemble(p, nmembers, forced, h) bind(C, name='bastr_create_ensemble')
      import
      type(bastr_params), intent(in) :: p
      integer(c_int32_t), value :: nmembers
      integer(c_int32_t), intent(in), optional :: forced(*)
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function bastr_nmembers(h) bind(C, name='bastr_nmembers')
      import
      type(c_ptr), value :: h
    end function
    integer(c_int) function bastr_set_member_fields(h, m, u1, u2, u3) bind(C, name='bastr_set_member_fields')
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: m
      real(c_double), intent(in) :: u1(*), u2(*)
      real(c_double), intent(in), optional :: u3(*)
    end function
    integer(c_int) function bastr_get_member_fields(h, m, u1, u2, u3) bind(C, name='bastr_get_member_fields')
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: m
      real(c_double), intent(out) :: u1(*), u2(*)
      real(c_double), intent(out), optional :: u3(*)
    end function
    integer(c_int) function bastr_init_random_phase_member(h, m, k0, seed, energy) &
        bind(C, name='bastr_init_random_phase_member')
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: m
      real(c_double), value :: k0, energy
      integer(c_int64_t), value :: seed
    end function
    integer(c_int) function bastr_get_member_time(h, m, nstep, time) bind(C, name='bastr_get_member_time')
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: m
      integer(c_int64_t), intent(out) :: nstep
      real(c_double), intent(out) :: time
    end function
    ! bastr_step on hs(1:nh) at once; diags(i) / spectra(i): c_loc of handle i's
    ! arrays (or c_null_ptr); the arrays themselves may be absent
    integer(c_int) function bastr_step_noalloc(h, n, diags, spectra) bind(C, name='bastr_step_noalloc')
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      type(bastr_diag), intent(out) :: diags(*)
      real(c_double), intent(out), optional :: spectra(*)
    end function
    integer(c_int) function bastr_step_group(hs, nh, n, diags, spectra) bind(C, name='bastr_step_group')
      import
      type(c_ptr), intent(in) :: hs(*)
      integer(c_int32_t), value :: nh, n
      type(c_ptr), intent(in), optional :: diags(*), spectra(*)
    end function
  end interface

contains

  ! The last error message of the calling thread as a Fortran string.
  function bastr_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: n
    interface
      integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
        import
        type(c_ptr), value :: s
      end function
    end interface
    type(c_ptr) :: p
    p = bastr_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, cs, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function

end module bastr_hip
