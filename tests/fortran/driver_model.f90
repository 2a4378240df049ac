! driver_model.f90 -- test program of the model-context part of the Fortran
! drop-in (exoplasim_amd/fortran/exo_legini_model.f90): built against the
! reference's own pumamod (src/plasimmod.f90 at src/resmod_def.f90, T21 L10,
! by oracle/Makefile), it sets the pumamod spectral state and calls legini
! (with a physics filter), invlega, invlegd, sp3fc and sp2fl by their
! reference names (src/legmod.f90:30,294,613,635,661), then writes the
! pumamod grid arrays they fill as real(8) to a stream file that
! tests/test_gpu_fortran.py compares with the oracle.
!   usage: driver_model NFILTER NQSPEC outfile
program exo_driver_model
  use pumamod
  implicit none
  integer :: kf, kq, i, k, l
  character(len=256) :: arg, outfile
  real :: zfc(NHOR, NLEV)

  call get_command_argument(1, arg); read(arg, *) kf
  call get_command_argument(2, arg); read(arg, *) kq
  call get_command_argument(3, outfile)

  ! prolog order of the model (src/plasim.f90:183,197-198,411)
  call inigau(NLAT, sid, gwd)
  nfilter = kf
  ngptfilter = 1
  nspvfilter = 1
  call legini
  nqspec = kq

  ! spectral state: golden(i, salt) per array, damped with (n+1)^-1.5
  do l = 1, NLEV
    do i = 1, NESP
      k = (l - 1) * NESP + i - 1
      sd(i, l) = golden(k, 0.3_8) * damp_of(i)
      st(i, l) = golden(k, 0.8_8) * damp_of(i)
      sz(i, l) = golden(k, 1.3_8) * damp_of(i)
      sq(i, l) = golden(k, 1.8_8) * damp_of(i)
    enddo
  enddo
  do i = 1, NESP
    sp(i) = golden(i - 1, 2.3_8) * damp_of(i)
  enddo

  open(10, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(10) real(sd, 8), real(st, 8), real(sz, 8), real(sq, 8), real(sp, 8)
  call clear
  call invlega
  write(10) real(gd, 8), real(gt, 8), real(gz, 8), real(gq, 8), real(gu, 8), real(gv, 8), &
            real(gp, 8), real(gpj, 8)
  call clear
  call invlegd
  write(10) real(gt, 8), real(gq, 8), real(gu, 8), real(gv, 8), real(gp, 8)
  call clear
  call sp3fc
  write(10) real(gd, 8), real(gt, 8), real(gz, 8)
  zfc(:, :) = 0.0
  call sp2fl(sq, zfc, NLEV)
  write(10) real(zfc, 8)
  close(10)
  write(*, '(a)') 'driver_model done'

contains
  subroutine clear
    gd = 0.0; gt = 0.0; gz = 0.0; gq = 0.0; gu = 0.0; gv = 0.0; gp = 0.0; gpj = 0.0
  end subroutine

  real(kind=8) function golden(i, s)
    integer, intent(in) :: i
    real(kind=8), intent(in) :: s
    golden = sin(0.7_8 * i + s) * cos(0.013_8 * real(i, 8)**1.5_8 + 2.0_8 * s)
  end function

  ! (n+1)^-1.5 of the spectral real i (1-based), Im(m=0) zeroed
  real(kind=8) function damp_of(i)
    integer, intent(in) :: i
    integer :: w, m, n, j
    w = (i + 1) / 2
    j = 0
    do m = 0, NTRU
      do n = m, NTRU
        j = j + 1
        if (j == w) then
          damp_of = (n + 1.0_8)**(-1.5_8)
          if (m == 0 .and. mod(i, 2) == 0) damp_of = 0.0_8
          return
        endif
      enddo
    enddo
    damp_of = 0.0_8
  end function
end program exo_driver_model

! the model's abort routine (src/mpimod.f90:372-386) as this test program
! has it: report the message and stop with a failure status
subroutine mpabort(ym)
  implicit none
  character(len=*) :: ym
  write(*, '(a)') 'Program aborted: ' // trim(ym)
  error stop 2
end subroutine mpabort
