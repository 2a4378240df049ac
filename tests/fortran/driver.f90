! driver.f90 -- test program of the Fortran drop-in (exoplasim_amd/fortran):
! calls the reference-named subroutines (inigau, sp2fc, sp2fcdmu, fc2sp,
! dv2uv, uv2dv, qtend, mktend, fc2gp, gp2fc) on deterministic inputs and
! writes its inputs and every result as real(8) to a stream file that
! tests/test_gpu_fortran.py compares with the oracle.
!   usage: driver NTRU NLAT NLEV outfile
program exo_driver
  use iso_c_binding
  use exo_legmod_capi
  implicit none
  integer :: ntru, nlat, nlev, nlon, ncsp, nfl, nhor, i, k, l, m, n, w
  character(len=256) :: arg, outfile
  real(kind=8), allocatable, target :: sid(:), gwd(:)
  real, allocatable, target :: sp(:,:), fc0(:), fcd(:), sp0(:), pu(:), pv(:)
  real, allocatable, target :: pd(:), pz(:), q(:), d(:), t(:), z(:), g(:), f(:,:)
  real(kind=8) :: salt

  call get_command_argument(1, arg); read(arg, *) ntru
  call get_command_argument(2, arg); read(arg, *) nlat
  call get_command_argument(3, arg); read(arg, *) nlev
  call get_command_argument(4, outfile)
  nlon = 2 * nlat
  ncsp = (ntru + 1) * (ntru + 2) / 2
  nfl = 3 * nlev + 1
  nhor = nlat * nlon
  allocate(sid(nlat), gwd(nlat), sp(2 * ncsp, nfl), fc0(nhor), fcd(nhor), sp0(2 * ncsp))
  allocate(pu(nhor * nlev), pv(nhor * nlev), pd(2 * ncsp * nlev), pz(2 * ncsp * nlev))
  allocate(q(2 * ncsp * nlev), d(2 * ncsp * nlev), t(2 * ncsp * nlev), z(2 * ncsp * nlev))
  allocate(g(nhor), f(nhor * nlev, 6))

  ! spectral input: golden_input of shape (nfl, ncsp, 2), (n+1)^-1.5, Im(m=0) = 0
  salt = 0.3_8
  do k = 1, nfl
    w = 0
    do m = 0, ntru
      do n = m, ntru
        w = w + 1
        do l = 1, 2
          i = ((k - 1) * ncsp + (w - 1)) * 2 + (l - 1)
          sp(2 * (w - 1) + l, k) = golden(i, salt) * (n + 1.0_8)**(-1.5_8)
          if (m == 0 .and. l == 2) sp(2 * (w - 1) + l, k) = 0.0
        enddo
      enddo
    enddo
  enddo
  do k = 1, 6
    do i = 0, nhor * nlev - 1
      f(i + 1, k) = golden(i, 1.1_8 + 0.5_8 * (k - 1))
    enddo
  enddo

  call inigau(nlat, sid, gwd)
  call exo_check(exo_legini(int(ntru, c_int), int(nlat, c_int), int(nlev, c_int), &
                 c_loc(sid(1)), c_loc(gwd(1)), c_null_ptr, c_null_ptr, 1.63299310207_c_double), &
                 'legini')
  call sp2fc(sp(1, 1), fc0)
  call sp2fcdmu(sp(1, nfl), fcd)
  call fc2sp(fc0, sp0)
  call dv2uv(sp(1, nlev + 1), sp(1, 1), pu, pv)
  call uv2dv(f(1, 1), f(1, 2), pd, pz)
  call qtend(q, f(1, 1), f(1, 2), f(1, 3))
  call mktend(d, t, z, f(1, 1), f(1, 2), f(1, 3), f(1, 4), f(1, 5), f(1, 6))
  g(:) = fc0(:)
  call fc2gp(g, nlon, nlat)

  open(10, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(10) real(sp, 8), real(f, 8)
  write(10) sid, gwd
  write(10) real(fc0, 8), real(fcd, 8), real(sp0, 8), real(pu, 8), real(pv, 8)
  write(10) real(pd, 8), real(pz, 8), real(q, 8), real(d, 8), real(t, 8), real(z, 8)
  write(10) real(g, 8)
  call gp2fc(g, nlon, nlat)
  write(10) real(g, 8)
  close(10)
  write(*, '(a)') 'driver done'

contains
  real(kind=8) function golden(i, s)
    integer, intent(in) :: i
    real(kind=8), intent(in) :: s
    golden = sin(0.7_8 * i + s) * cos(0.013_8 * real(i, 8)**1.5_8 + 2.0_8 * s)
  end function
end program exo_driver

! error path of the test program: report and stop with a failure status
subroutine exo_abort(msg)
  implicit none
  character(len=*) :: msg
  write(*, '(a)') trim(msg)
  error stop 2
end subroutine exo_abort
