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
