! the abort path of the drop-in for test programs and libraries that run
! without the model: print the message and stop
subroutine exo_abort(msg)
  implicit none
  character(len=*) :: msg
  write(*, '(a)') trim(msg)
  stop 2
end subroutine exo_abort
