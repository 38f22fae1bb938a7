! Drives the drop-in fftwlink module the way Bastr.F90 does
! (fftw_mpi_init, mpisizedis_fftw, allocation, prepareplan_fftw_plan,
! fft2d/3d, ifft2d/3d, deallocation) on a fixed complex field, and writes the
! forward transform to fftwlink_fwd.bin for the test to compare with numpy.
! usage: fftwlink_check ia ja ka     (ka = 0: 2-D)
program fftwlink_check
  use, intrinsic :: iso_c_binding
  use commvar
  use fftwlink
  use bastr_single_rank
  implicit none
  interface
    subroutine fftw_mpi_init() bind(C, name='fftw_mpi_init')
    end subroutine
  end interface
  complex(c_double_complex), allocatable :: ref(:,:,:)
  integer :: i, j, k, nz
  character(len=32) :: arg
  real(8) :: err

  call get_command_argument(1, arg); read(arg, *) ia
  call get_command_argument(2, arg); read(arg, *) ja
  call get_command_argument(3, arg); read(arg, *) ka
  ndims = merge(3, 2, ka > 0)
  timemethod = 1
  allkmax = ia / 3
  call fftw_mpi_init()
  call mpisizedis_fftw
  call allocation
  call prepareplan_fftw_plan
  nz = size(u1spe, 3)
  write(*, '(A,4I8)') ' sizes', size(u1spe, 1), size(u1spe, 2), nz, int(alloc_local)
  do k = 1, nz
    do j = 1, ja
      do i = 1, ia
        u1spe(i, j, k) = cmplx(cos(0.3d0 * i + 0.7d0 * j + 1.1d0 * k) + 0.01d0 * i, &
                               sin(0.05d0 * i * j) - 0.02d0 * k, c_double_complex)
      end do
    end do
  end do
  allocate(ref, source=u1spe)
  if (ndims == 2) then
    call fft2d(u1spe)
  else
    call fft3d(u1spe)
  end if
  open(10, file='fftwlink_fwd.bin', access='stream', form='unformatted', status='replace')
  write(10) u1spe
  close(10)
  if (ndims == 2) then
    call ifft2d(u1spe)
  else
    call ifft3d(u1spe)
  end if
  err = maxval(abs(u1spe - ref)) / maxval(abs(ref))
  write(*, '(A,ES12.4)') ' roundtrip', err
  write(*, '(A,2F8.3)') ' reductions', psum(1.5d0), pmin(2.5d0)
  call deallocation
end program fftwlink_check
