! Test fixture: the end-of-run routine the drop-in fftwlink's deallocation
! calls (one rank, no MPI).
module parallel
  implicit none
contains
  subroutine mpistop
    print *, ' ** The job is done!'
    stop
  end subroutine
end module parallel
