! exo_abort_model.f90 -- error path of the drop-in inside the model: the
! model's mpabort writes Abort_Message and stops (src/mpimod.f90:330-390).
subroutine exo_abort(msg)
  implicit none
  character(len=*) :: msg
  call mpabort(msg)
end subroutine exo_abort
