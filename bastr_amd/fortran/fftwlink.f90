! fftwlink.f90 -- drop-in replacement for the reference's fftwlink module
! (src/fftwlink.F90): the same public names and argument shapes, one rank
! holding the whole domain, and the transforms on the GPU through the
! bastr_hip binding (bastr_fft_forward / bastr_fft_backward, which keep the
! forward x 1/(ia*ja[*ka]), backward unnormalised convention).
!
! Built in place of src/fftwlink.F90 against the program's own commvar and
! parallel modules, an unmodified Bastr.F90 / Bastrpp.F90 links without
! FFTW: this file also defines the C symbol fftw_mpi_init (a no-op), the
! one FFTW routine those programs call directly.  Buffers are Fortran
! allocations; the "plans" are one bastr handle per run.  Grids are those of
! bastr_create (2-D powers of two in [16, 4096]; 3-D cubic, [16, 512]).
!
! For speed, a caller replaces mainloop's body with bastr_step (bastr_hip);
! these transforms copy each field to the GPU and back.
module fftwlink
  use, intrinsic :: iso_c_binding
  use commvar, only: im, jm, km, ia, ja, ka, ndims
  use bastr_hip
  implicit none

  integer(C_INTPTR_T) :: alloc_local, iafftw, jafftw, kafftw, imfftw, jmfftw, kmfftw
  integer :: nproc = 1, myid = 0, ierr = 0
  integer :: i0f = 0, j0f = 0, k0f = 0
  integer :: rooti0 = 0, imf, jmf, kmf
  integer, allocatable, dimension(:) :: counts_grid, disp_grid, counts_fence, disp_fence
  ! both "plans" hold the one bastr handle that runs the transforms
  type(C_PTR) :: forward_plan = C_NULL_PTR, backward_plan = C_NULL_PTR

contains

  ! fftw_mpi_init of Bastr.F90:26 / Bastrpp.F90:26: nothing to set up on one rank
  subroutine bastr_fftw_mpi_init() bind(C, name='fftw_mpi_init')
  end subroutine

  subroutine report_sizes()
    write(*, '(3(A,I0))') '  ** mpi size= ', 1, ' x ', 1, ' x ', 1
  end subroutine

  ! mpisizedis_fftw (fftwlink.F90:34-153) with one rank: the local block is
  ! the whole grid, offsets zero; sets commvar's im, jm, km.
  subroutine mpisizedis_fftw
    iafftw = ia
    jafftw = ja
    kafftw = ka
    imfftw = ia
    jmfftw = ja
    select case (ndims)
    case (3)
      kmfftw = ka
      alloc_local = int(ia, C_INTPTR_T) * ja * ka
    case (2)
      kmfftw = 0
      alloc_local = int(ia, C_INTPTR_T) * ja
    case default
      stop 'fftw_mpi: 1D not implemented!'
    end select
    imf = int(imfftw)
    jmf = int(jmfftw)
    kmf = int(kmfftw)
    im = imf
    jm = jmf
    km = kmf
    i0f = 0
    j0f = 0
    k0f = 0
    call report_sizes()
  end subroutine mpisizedis_fftw

  ! mpisizedis_half_fftw (fftwlink.F90:158-280): the r2c half-spectrum sizes
  subroutine mpisizedis_half_fftw
    iafftw = ia
    jafftw = ja
    kafftw = ka
    imfftw = ia / 2 + 1
    jmfftw = ja
    select case (ndims)
    case (3)
      kmfftw = ka
      alloc_local = imfftw * ja * ka
    case (2)
      kmfftw = 0
      alloc_local = imfftw * ja
    case default
      stop 'fftw_mpi: 1D not implemented!'
    end select
    imf = int(imfftw)
    jmf = int(jmfftw)
    kmf = int(kmfftw)
    im = imf
    jm = jmf
    km = kmf
    i0f = 0
    j0f = 0
    k0f = 0
    call report_sizes()
  end subroutine mpisizedis_half_fftw

  ! prepareplan_fftw_plan (fftwlink.F90:284-303): one FP64 handle for the grid
  subroutine prepareplan_fftw_plan
    type(bastr_params) :: p
    type(c_ptr) :: h
    if (ndims /= 2 .and. ndims /= 3) stop 'prepareplan_fftw: Not implemented for 1D!'
    p = bastr_params(ia, ja, merge(ka, 0, ndims == 3), 1.d0, 1.d0, 1, 0, 5, 1.d0, 0, 1, &
                     BASTR_PREC_F64)
    if (bastr_create(p, h) /= BASTR_OK) then
      write(*, '(A,A)') ' !! ', bastr_error()
      stop 'prepareplan_fftw: grid not supported'
    end if
    forward_plan = h
    backward_plan = h
  end subroutine prepareplan_fftw_plan

  ! allocate_fftw_complex (fftwlink.F90:305-322): an (imfftw, jmfftw, 1 | kmfftw)
  ! buffer and its address
  subroutine allocate_fftw_complex(array, c_pointer)
    type(C_PTR), intent(out) :: c_pointer
    complex(C_DOUBLE_COMPLEX), pointer, intent(out) :: array(:,:,:)
    select case (ndims)
    case (2)
      allocate(array(imfftw, jmfftw, 1))
    case (3)
      allocate(array(imfftw, jmfftw, kmfftw))
    case default
      stop 'allocate_fftw_complex: Not implemented for 1D!'
    end select
    c_pointer = c_loc(array(1, 1, 1))
  end subroutine allocate_fftw_complex

  subroutine run_transform(array, forward)
    complex(C_DOUBLE_COMPLEX), pointer, intent(inout) :: array(:,:,:)
    logical, intent(in) :: forward
    integer(c_int) :: rc
    if (.not. c_associated(forward_plan)) stop 'fftwlink: prepareplan_fftw_plan not called'
    if (forward) then
      rc = bastr_fft_forward(forward_plan, array)
    else
      rc = bastr_fft_backward(backward_plan, array)
    end if
    if (rc /= BASTR_OK) then
      write(*, '(A,A)') ' !! ', bastr_error()
      stop 'fftwlink: transform failed'
    end if
  end subroutine run_transform

  ! fft2d / fft3d (fftwlink.F90:324-348): forward, divided by ia*ja[*ka]
  subroutine fft2d(array)
    complex(C_DOUBLE_COMPLEX), pointer, intent(inout) :: array(:,:,:)
    call run_transform(array, .true.)
  end subroutine fft2d

  subroutine fft3d(array)
    complex(C_DOUBLE_COMPLEX), pointer, intent(inout) :: array(:,:,:)
    call run_transform(array, .true.)
  end subroutine fft3d

  ! ifft2d / ifft3d (fftwlink.F90:350-368): backward, unnormalised
  subroutine ifft2d(array)
    complex(C_DOUBLE_COMPLEX), pointer, intent(inout) :: array(:,:,:)
    call run_transform(array, .false.)
  end subroutine ifft2d

  subroutine ifft3d(array)
    complex(C_DOUBLE_COMPLEX), pointer, intent(inout) :: array(:,:,:)
    call run_transform(array, .false.)
  end subroutine ifft3d

  ! allocation (fftwlink.F90:370-467): the program's fields and work buffers
  subroutine allocation
    use commvar
    select case (ndims)
    case (2)
      call allocation2D
    case (3)
      call allocation3D
    case default
      stop 'allocation: Not implemented for 1D!'
    end select
    allocate(Es(0:allkmax), Ed(0:allkmax), kn(0:allkmax), Ecount(0:allkmax))
  end subroutine allocation

  subroutine allocation2D
    use commvar
    allocate(u1(1:im, 1:jm, 0:km), u2(1:im, 1:jm, 0:km))
    allocate(k1(1:im, 1:jm, 0:km), k2(1:im, 1:jm, 0:km))
    allocate(random_angle(1:im, 1:jm, 0:km))
    call allocate_fftw_complex(u1spe, c_u1spe)
    call allocate_fftw_complex(u2spe, c_u2spe)
    call allocate_fftw_complex(u1x1, c_u1x1)
    call allocate_fftw_complex(u1x2, c_u1x2)
    call allocate_fftw_complex(u2x1, c_u2x1)
    call allocate_fftw_complex(u2x2, c_u2x2)
    call allocate_fftw_complex(u1xixi, c_u1xixi)
    call allocate_fftw_complex(u2xixi, c_u2xixi)
    call allocate_fftw_complex(thetaxixi, c_thetaxixi)
    call allocate_fftw_complex(u1tA, c_u1tA)
    call allocate_fftw_complex(u2tA, c_u2tA)
    call allocate_fftw_complex(force1, c_force1)
    call allocate_fftw_complex(force2, c_force2)
    call allocate_fftw_complex(random_complex, c_randomcomplex)
    select case (timemethod)
    case (1)
      call allocate_fftw_complex(u1tB, c_u1tB)
      call allocate_fftw_complex(u2tB, c_u2tB)
      call allocate_fftw_complex(u1tC, c_u1tC)
      call allocate_fftw_complex(u2tC, c_u2tC)
    case (2)
      allocate(u1old(1:im, 1:jm, 0:km), u2old(1:im, 1:jm, 0:km))
    case default
      stop 'allocation2D: timemethod not recognized!'
    end select
  end subroutine allocation2D

  subroutine allocation3D
    use commvar
    allocate(u1(1:im, 1:jm, 1:km), u2(1:im, 1:jm, 1:km), u3(1:im, 1:jm, 1:km))
    allocate(k1(1:im, 1:jm, 1:km), k2(1:im, 1:jm, 1:km), k3(1:im, 1:jm, 1:km))
    allocate(random_angle(1:im, 1:jm, 1:km))
    call allocate_fftw_complex(u1spe, c_u1spe)
    call allocate_fftw_complex(u2spe, c_u2spe)
    call allocate_fftw_complex(u3spe, c_u3spe)
    call allocate_fftw_complex(u1x1, c_u1x1)
    call allocate_fftw_complex(u1x2, c_u1x2)
    call allocate_fftw_complex(u1x3, c_u1x3)
    call allocate_fftw_complex(u2x1, c_u2x1)
    call allocate_fftw_complex(u2x2, c_u2x2)
    call allocate_fftw_complex(u2x3, c_u2x3)
    call allocate_fftw_complex(u3x1, c_u3x1)
    call allocate_fftw_complex(u3x2, c_u3x2)
    call allocate_fftw_complex(u3x3, c_u3x3)
    call allocate_fftw_complex(u1xixi, c_u1xixi)
    call allocate_fftw_complex(u2xixi, c_u2xixi)
    call allocate_fftw_complex(u3xixi, c_u3xixi)
    call allocate_fftw_complex(thetaxixi, c_thetaxixi)
    call allocate_fftw_complex(u1tA, c_u1tA)
    call allocate_fftw_complex(u2tA, c_u2tA)
    call allocate_fftw_complex(u3tA, c_u3tA)
    call allocate_fftw_complex(force1, c_force1)
    call allocate_fftw_complex(force2, c_force2)
    call allocate_fftw_complex(force3, c_force3)
    call allocate_fftw_complex(random_complex, c_randomcomplex)
    select case (timemethod)
    case (1)
      call allocate_fftw_complex(u1tB, c_u1tB)
      call allocate_fftw_complex(u2tB, c_u2tB)
      call allocate_fftw_complex(u3tB, c_u3tB)
      call allocate_fftw_complex(u1tC, c_u1tC)
      call allocate_fftw_complex(u2tC, c_u2tC)
      call allocate_fftw_complex(u3tC, c_u3tC)
    case (2)
      allocate(u1old(1:im, 1:jm, 1:km), u2old(1:im, 1:jm, 1:km), u3old(1:im, 1:jm, 1:km))
    case default
      stop 'allocation3D: timemethod not recognized!'
    end select
  end subroutine allocation3D

  subroutine free_buffer(array, c_pointer)
    complex(C_DOUBLE_COMPLEX), pointer, intent(inout) :: array(:,:,:)
    type(C_PTR), intent(inout) :: c_pointer
    if (associated(array)) deallocate(array)
    c_pointer = C_NULL_PTR
  end subroutine free_buffer

  ! deallocation (fftwlink.F90:469-562): frees everything, releases the
  ! handle and ends the run through the program's mpistop, as the reference does
  subroutine deallocation
    use commvar
    use parallel, only: mpistop
    deallocate(Es, Ed, kn, Ecount)
    if (c_associated(forward_plan)) then
      if (bastr_destroy(forward_plan) /= BASTR_OK) write(*, '(A,A)') ' !! ', bastr_error()
    end if
    forward_plan = C_NULL_PTR
    backward_plan = C_NULL_PTR
    select case (ndims)
    case (3)
      call deallocation3D
    case (2)
      call deallocation2D
    end select
    call mpistop
  end subroutine deallocation

  subroutine deallocation2D
    use commvar
    deallocate(u1, u2, k1, k2, random_angle)
    call free_buffer(u1spe, c_u1spe)
    call free_buffer(u2spe, c_u2spe)
    call free_buffer(u1x1, c_u1x1)
    call free_buffer(u1x2, c_u1x2)
    call free_buffer(u2x1, c_u2x1)
    call free_buffer(u2x2, c_u2x2)
    call free_buffer(u1xixi, c_u1xixi)
    call free_buffer(u2xixi, c_u2xixi)
    call free_buffer(thetaxixi, c_thetaxixi)
    call free_buffer(u1tA, c_u1tA)
    call free_buffer(u2tA, c_u2tA)
    call free_buffer(force1, c_force1)
    call free_buffer(force2, c_force2)
    call free_buffer(random_complex, c_randomcomplex)
    if (timemethod == 1) then
      call free_buffer(u1tB, c_u1tB)
      call free_buffer(u2tB, c_u2tB)
      call free_buffer(u1tC, c_u1tC)
      call free_buffer(u2tC, c_u2tC)
    else if (timemethod == 2) then
      deallocate(u1old, u2old)
    end if
  end subroutine deallocation2D

  subroutine deallocation3D
    use commvar
    deallocate(u1, u2, u3, k1, k2, k3, random_angle)
    call free_buffer(u1spe, c_u1spe)
    call free_buffer(u2spe, c_u2spe)
    call free_buffer(u3spe, c_u3spe)
    call free_buffer(u1x1, c_u1x1)
    call free_buffer(u1x2, c_u1x2)
    call free_buffer(u1x3, c_u1x3)
    call free_buffer(u2x1, c_u2x1)
    call free_buffer(u2x2, c_u2x2)
    call free_buffer(u2x3, c_u2x3)
    call free_buffer(u3x1, c_u3x1)
    call free_buffer(u3x2, c_u3x2)
    call free_buffer(u3x3, c_u3x3)
    call free_buffer(u1xixi, c_u1xixi)
    call free_buffer(u2xixi, c_u2xixi)
    call free_buffer(u3xixi, c_u3xixi)
    call free_buffer(thetaxixi, c_thetaxixi)
    call free_buffer(u1tA, c_u1tA)
    call free_buffer(u2tA, c_u2tA)
    call free_buffer(u3tA, c_u3tA)
    call free_buffer(force1, c_force1)
    call free_buffer(force2, c_force2)
    call free_buffer(force3, c_force3)
    call free_buffer(random_complex, c_randomcomplex)
    if (timemethod == 1) then
      call free_buffer(u1tB, c_u1tB)
      call free_buffer(u2tB, c_u2tB)
      call free_buffer(u3tB, c_u3tB)
      call free_buffer(u1tC, c_u1tC)
      call free_buffer(u2tC, c_u2tC)
      call free_buffer(u3tC, c_u3tC)
    else if (timemethod == 2) then
      deallocate(u1old, u2old, u3old)
    end if
  end subroutine deallocation3D

end module fftwlink
