! bastr_single_rank.f90 -- one-rank reductions with the generic names of the
! reference's psum / pmin / pmax (parallel.F90:2321-2482, 2616-2640), for a
! program built around the GPU step without MPI: one rank holds the whole
! field, so every reduction returns its argument.
module bastr_single_rank
  implicit none
  private
  public :: psum, pmin, pmax

  interface psum
    module procedure sum_i4, sum_r8, sum_i4_1d, sum_r8_1d, sum_r8_2d, sum_r8_3d
  end interface
  interface pmin
    module procedure ext_i4, ext_r8
  end interface
  interface pmax
    module procedure ext_i4, ext_r8
  end interface

contains

  integer function sum_i4(v)
    integer, intent(in) :: v
    sum_i4 = v
  end function
  real(8) function sum_r8(v)
    real(8), intent(in) :: v
    sum_r8 = v
  end function
  function sum_i4_1d(v) result(r)
    integer, intent(in) :: v(:)
    integer :: r(size(v))
    r = v
  end function
  function sum_r8_1d(v) result(r)
    real(8), intent(in) :: v(:)
    real(8) :: r(size(v))
    r = v
  end function
  function sum_r8_2d(v) result(r)
    real(8), intent(in) :: v(:,:)
    real(8) :: r(size(v, 1), size(v, 2))
    r = v
  end function
  function sum_r8_3d(v) result(r)
    real(8), intent(in) :: v(:,:,:)
    real(8) :: r(size(v, 1), size(v, 2), size(v, 3))
    r = v
  end function
  integer function ext_i4(v)
    integer, intent(in) :: v
    ext_i4 = v
  end function
  real(8) function ext_r8(v)
    real(8), intent(in) :: v
    ext_r8 = v
  end function

end module bastr_single_rank
