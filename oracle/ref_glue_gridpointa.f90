! ref_glue_gridpointa.f90 -- TEST INFRASTRUCTURE ONLY (part of the oracle).
!
! C-callable access to the reference's gridpointa (src/plasim.f90:2360-2585)
! at its default resolution module (src/resmod_def.f90: T21 L10, NPRO=1),
! built with legmod, gaussmod and fftmod so that the whole transform group
! runs: invlega, gpmt, fc2gp, exp, calcgp, the products, gp2fc, fc2sp,
! mktend, qtend and mpsumsc.  The diagnostic branches are switched off
! (ngui = 0, nstep not a multiple of ndiag or nafter, nentropy = nenergy = 0).
! Built only by oracle/Makefile into oracle/_ref/ and loaded with lazy
! binding: plasim.f90 references model subroutines this library does not
! link, none of which this call reaches.

! Spectral state (NESP, 4 NLEV + 1) = sz sd st sq sp in, tendencies in the
! same layout out (sqt zero when kq = 0); pgq and pdq the grid-point
! humidity the model holds when nqspec = 0; the vertical set-up as
! ref_plasim_calcgp takes it; pgrid (NHOR, 6 NLEV + 2) out: the pumamod grid
! fields gridpointa leaves behind, gz gd gt gu gv (after fc2gp) and dw, then
! gp (= exp(ln ps)) and gpj.  Its other fields are locals of gridpointa.
subroutine ref_gridpointa(pst, psigmah, pdsigma, prdsig, pt0, pakap, pscal, kq, pgq, pdq, &
                          ptend, pgrid) bind(C, name="ref_gridpointa")
  use pumamod
  real(8) :: pst(NESP, 4 * NLEV + 1), psigmah(NLEV), pdsigma(NLEV), prdsig(NLEV)
  real(8) :: pt0(NLEV), pakap, pscal(4), pgq(NHOR, NLEV), pdq(NHOR, NLEV)
  integer(4), value :: kq
  real(8) :: ptend(NESP, 4 * NLEV + 1), pgrid(NHOR, 6 * NLEV + 2)
  integer :: jlat, n
  n = NLEV
  call inigau(NLAT, sid, gwd)
  do jlat = 1, NLAT
    csq(jlat) = 1.0 - sid(jlat) * sid(jlat)
    rcsq(1 + (jlat - 1) * NLON:jlat * NLON) = 1.0 / csq(jlat)
  enddo
  call legini
  sigmah(:) = psigmah(:)
  dsigma(:) = pdsigma(:)
  rdsig(:) = prdsig(:)
  t0(:) = pt0(:)
  akap = pakap
  rdbrv = pscal(1)
  adv = pscal(2)
  psurf = pscal(3)
  ww = pscal(4)
  nqspec = kq
  nprint = 0
  call initsi
  sz(:, :) = pst(:, 1:n)
  sd(:, :) = pst(:, n + 1:2 * n)
  st(:, :) = pst(:, 2 * n + 1:3 * n)
  sq(:, :) = pst(:, 3 * n + 1:4 * n)
  sp(:) = pst(:, 4 * n + 1)
  gq(:, :) = pgq(:, :)
  dq(:, 1:NLEV) = pdq(:, :)
  sdt(:, :) = 0.0
  stt(:, :) = 0.0
  szt(:, :) = 0.0
  sqt(:, :) = 0.0
  spt(:) = 0.0
  ngui = 0
  ndiag = 1000
  nafter = 1000
  nstep = 1
  nentropy = 0
  nenergy = 0
  call gridpointa
  ptend(:, 1:n) = szt(:, :)
  ptend(:, n + 1:2 * n) = sdt(:, :)
  ptend(:, 2 * n + 1:3 * n) = stt(:, :)
  ptend(:, 3 * n + 1:4 * n) = sqt(:, :)
  ptend(:, 4 * n + 1) = spt(:)
  pgrid(:, 1:n) = gz(:, :)
  pgrid(:, n + 1:2 * n) = gd(:, :)
  pgrid(:, 2 * n + 1:3 * n) = gt(:, :)
  pgrid(:, 3 * n + 1:4 * n) = gu(:, :)
  pgrid(:, 4 * n + 1:5 * n) = gv(:, :)
  pgrid(:, 5 * n + 1:6 * n) = dw(:, :)
  pgrid(:, 6 * n + 1) = gp(:)
  pgrid(:, 6 * n + 2) = gpj(:)
end subroutine
