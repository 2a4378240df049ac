! ref_glue_t21.f90 -- TEST INFRASTRUCTURE ONLY (part of the oracle).
!
! C-callable set-up and table access for the reference legmod built at the
! reference's own default resolution module (src/resmod_def.f90: NLAT=32,
! NLEV=10, NPRO=1, i.e. T21 L10).  The transforms themselves (sp2fc_,
! fc2sp_, dv2uv_, ... and fftmod's gp2fc_/fc2gp_) are called directly
! through their implicit-interface symbols.  Built only by oracle/Makefile
! into oracle/_ref/.

subroutine ref_t21_dims(d) bind(C, name="ref_t21_dims")
  use pumamod
  integer(4) :: d(6)
  d(1) = NLAT
  d(2) = NLON
  d(3) = NTRU
  d(4) = NLEV
  d(5) = NCSP
  d(6) = NESP
end subroutine

! inigau + legini exactly as prolog does it (src/plasim.f90:183,411)
subroutine ref_t21_init(psid, pgwd) bind(C, name="ref_t21_init")
  use pumamod
  real(8) :: psid(NLAT), pgwd(NLAT)
  call inigau(NLAT, sid, gwd)
  call legini
  psid(:) = sid(:)
  pgwd(:) = gwd(:)
end subroutine

! copy the eight legmod tables out as tab(NCSP, NLAT, 8) in the order
! qi, qj, qc, qe, qm, qq, qu, qv
subroutine ref_t21_tables(tab) bind(C, name="ref_t21_tables")
  use legmod
  real(8) :: tab(NCSP, NLPP, 8)
  tab(:,:,1) = qi
  tab(:,:,2) = qj
  tab(:,:,3) = qc
  tab(:,:,4) = qe
  tab(:,:,5) = qm
  tab(:,:,6) = qq
  tab(:,:,7) = qu
  tab(:,:,8) = qv
end subroutine

! plavor as the model holds it (src/plasimmod.f90:227)
subroutine ref_t21_plavor(p) bind(C, name="ref_t21_plavor")
  use pumamod
  real(8) :: p
  p = plavor
end subroutine

! one member round trip with the reference's own subroutines, in the order
! of the oracle's ora_roundtrip: sp2fc of the nfl field-levels, sp2fcdmu of
! ln ps, dv2uv of the NLEV levels, fc2gp of everything, then gp2fc + fc2sp
! of the nfl prognostic fields (used to time the oracle against the
! reference on the same CPU)
subroutine ref_t21_roundtrip(pspin, pgp, pspout) bind(C, name="ref_t21_roundtrip")
  use pumamod
  integer, parameter :: nfl = 3 * NLEV + 1, ngp = nfl + 1 + 2 * NLEV
  real(8) :: pspin(NESP, nfl), pgp(NHOR, ngp), pspout(NESP, nfl)
  real(8) :: zz(NESP, NLEV), zw(NHOR)
  integer :: jf
  do jf = 1, nfl
    call sp2fc(pspin(1, jf), pgp(1, jf))
  enddo
  call sp2fcdmu(pspin(1, nfl), pgp(1, nfl + 1))
  zz(:, :) = pspin(:, 1:NLEV)
  call dv2uv(pspin(1, NLEV + 1), zz, pgp(1, nfl + 2), pgp(1, nfl + 2 + NLEV))
  call fc2gp(pgp, NLON, NLAT * ngp)
  do jf = 1, nfl
    zw(:) = pgp(:, jf)
    call gp2fc(zw, NLON, NLAT)
    call fc2sp(zw, pspout(1, jf))
  enddo
end subroutine

! physics filter of legini (src/legmod.f90:123-155): set the namelist
! switches, then run legini again; the legmod filters out as
! skgpsp(NTRU+1), skspgp(NTRU+1)
subroutine ref_t21_set_filter(kfilter, kgpt, kspv, pkappa, kexp, pn0, pgpsp, pspgp) &
    bind(C, name="ref_t21_set_filter")
  use legmod
  integer(4), value :: kfilter, kgpt, kspv, kexp
  real(8), value :: pkappa, pn0
  real(8) :: pgpsp(NTP1), pspgp(NTP1)
  nfilter = kfilter
  ngptfilter = kgpt
  nspvfilter = kspv
  filterkappa = pkappa
  nfilterexp = kexp
  landhoskn0 = pn0
  call legini
  pgpsp(:) = skgpsp(:)
  pspgp(:) = skspgp(:)
end subroutine
