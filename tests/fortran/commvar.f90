! Test fixture: the program-state module that the drop-in fftwlink module
! reads and allocates into.  Declares the names fftwlink.f90 uses (grid
! sizes, fields, work buffers, shell arrays), nothing else.
module commvar
  use, intrinsic :: iso_c_binding
  implicit none
  integer :: ia, ja, ka, im, jm, km, ndims
  integer :: timemethod, allkmax
  real(8), allocatable, dimension(:,:,:) :: u1, u2, u3, u1old, u2old, u3old, k1, k2, k3, random_angle
  real(8), allocatable, dimension(:) :: Es, Ed, kn
  integer, allocatable, dimension(:) :: Ecount
  type(C_PTR) :: c_u1spe, c_u2spe, c_u3spe, c_u1x1, c_u1x2, c_u1x3, c_u2x1, c_u2x2, c_u2x3, &
                 c_u3x1, c_u3x2, c_u3x3, c_u1xixi, c_u2xixi, c_u3xixi, c_thetaxixi, &
                 c_u1tA, c_u2tA, c_u3tA, c_u1tB, c_u2tB, c_u3tB, c_u1tC, c_u2tC, c_u3tC, &
                 c_force1, c_force2, c_force3, c_randomcomplex
  complex(c_double_complex), pointer, dimension(:,:,:) :: u1spe, u2spe, u3spe, u1x1, u1x2, u1x3, &
      u2x1, u2x2, u2x3, u3x1, u3x2, u3x3, u1xixi, u2xixi, u3xixi, thetaxixi, &
      u1tA, u2tA, u3tA, u1tB, u2tB, u3tB, u1tC, u2tC, u3tC, force1, force2, force3, random_complex
end module commvar
