! exo_legini_model.f90 -- the pumamod-dependent part of the drop-in:
! legini, invlega, invlegd and sp3fc with the reference's names
! (src/legmod.f90:30, 294, 635, 661).  sp3fc, invlega and invlegd make the
! reference's calls in fewer of them: sp2fl transforms all levels of a field
! in one call where the reference loops over levels.  It is compiled inside the model build
! (it uses the model's module pumamod) together with exo_legmod_capi.f90,
! exo_transforms.f90 and exo_abort_model.f90.

! legini (src/legmod.f90:30): tables on the GPU from the model's sid/gwd
subroutine legini
  use pumamod
  use exo_legmod_capi
  implicit none
  real(kind=8), target :: zsid(NLAT), zgwd(NLAT), zkg(NTP1), zks(NTP1)
  integer :: n
  ! the filter is evaluated in the model's default real kind, as the
  ! reference's legini does, so both MOST_PREC builds get its values
  real :: f, g1, s1
  zsid(:) = sid(:)
  zgwd(:) = gwd(:)
  ! physics filters with 1-based n (src/legmod.f90:123-155)
  do n = 1, NTP1
    select case (nfilter)
    case (1)
      f = 1.0 - real(n) / (NTP1)
    case (2)
      f = exp(-filterkappa * (real(n) / NTRU)**nfilterexp)
    case (3)
      f = exp(-(real(n) * real(n + 1) / (landhoskn0 * (landhoskn0 + 1)))**2)
    case (4)
      f = (1.0 - real(n) / (NTP1))**2
    case default
      f = 1.0
    end select
    g1 = 1.0
    s1 = 1.0
    if (nfilter >= 1 .and. nfilter <= 4) then
      g1 = (1 - ngptfilter) * g1 + ngptfilter * f
      s1 = (1 - nspvfilter) * s1 + nspvfilter * f
    endif
    zkg(n) = real(g1, 8)
    zks(n) = real(s1, 8)
  enddo
  if (NPRO > 1) call exo_abort('exoplasim_amd: legini supports NPRO=1 only')
  call exo_check(exo_legini(int(NTRU, c_int), int(NLAT, c_int), int(NLEV, c_int), &
                 c_loc(zsid(1)), c_loc(zgwd(1)), c_loc(zkg(1)), c_loc(zks(1)), &
                 real(plavor, c_double)), 'legini')
end subroutine legini

! sp3fc (src/legmod.f90:294)
subroutine sp3fc
  use pumamod, only: NLEV, sd, st, sz, gd, gt, gz
  implicit none
  call sp2fl(sd, gd, NLEV)
  call sp2fl(st, gt, NLEV)
  call sp2fl(sz, gz, NLEV)
end subroutine sp3fc

! invlega (src/legmod.f90:635)
subroutine invlega
  use pumamod
  implicit none
  call dv2uv(sd, sz, gu, gv)
  call sp2fl(sd, gd, NLEV)
  call sp2fl(st, gt, NLEV)
  call sp2fl(sz, gz, NLEV)
  if (nqspec == 1) call sp2fl(sq, gq, NLEV)
  call sp2fc(sp, gp)
  call sp2fcdmu(sp, gpj)
end subroutine invlega

! invlegd (src/legmod.f90:661)
subroutine invlegd
  use pumamod
  implicit none
  call dv2uv(sd, sz, gu, gv)
  if (nqspec == 1) call sp2fl(sq, gq, NLEV)
  call sp2fl(st, gt, NLEV)
  call sp2fc(sp, gp)
end subroutine invlegd
