! ref_glue_plasim.f90 -- TEST INFRASTRUCTURE ONLY (part of the oracle).
!
! C-callable access to routines of the reference's plasim.f90 built at the
! reference's own default resolution module (src/resmod_def.f90: T21 L10,
! NPRO=1): the semi-implicit matrices (initsi, makebm, minvers), the
! semi-implicit part of spectrala (sections 1-3, src/plasim.f90:2817-2876)
! and the grid-point tendencies calcgp (src/plasim.f90:2587-2750).  Each
! routine copies its inputs into the pumamod globals the reference routine
! reads, calls it, and copies the globals it writes back out.  Built only by
! oracle/Makefile into oracle/_ref/ and loaded with lazy binding: plasim.f90
! references model subroutines this library does not link, none of which
! these calls reach.

subroutine ref_plasim_dims(d) bind(C, name="ref_plasim_dims")
  use pumamod
  integer(4) :: d(8)
  d(1) = NLAT
  d(2) = NLON
  d(3) = NTRU
  d(4) = NLEV
  d(5) = NCSP
  d(6) = NESP
  d(7) = NHOR
  d(8) = NSPP
end subroutine

! vertical coordinate and reference profile in, initsi's matrices out
! (src/plasim.f90:1723-1770): g, tau, c as (NLEV, NLEV) column-major
subroutine ref_plasim_initsi(psigma, psigmah, pdsigma, prdsig, pt0, pakap, &
                             pg, ptau, pc, pt01s2, ptkp) bind(C, name="ref_plasim_initsi")
  use pumamod
  real(8) :: psigma(NLEV), psigmah(NLEV), pdsigma(NLEV), prdsig(NLEV), pt0(NLEV)
  real(8) :: pakap
  real(8) :: pg(NLEV, NLEV), ptau(NLEV, NLEV), pc(NLEV, NLEV), pt01s2(NLEV), ptkp(NLEV)
  sigma(:) = psigma(:)
  sigmah(:) = psigmah(:)
  dsigma(:) = pdsigma(:)
  rdsig(:) = prdsig(:)
  t0(:) = pt0(:)
  akap = pakap
  nprint = 0
  call initsi
  pg(:, :) = g(:, :)
  ptau(:, :) = tau(:, :)
  pc(:, :) = c(:, :)
  pt01s2(:) = t01s2(:)
  ptkp(:) = tkp(:)
end subroutine

! makebm (src/plasim.f90:1683-1703) on the g, tau, t0, dsigma set above and
! the given matrices; bm1 out as (NLEV, NLEV, NTRU) column-major
subroutine ref_plasim_makebm(pdelt, pt0, pdsigma, pg, ptau, pbm1) &
    bind(C, name="ref_plasim_makebm")
  use pumamod
  real(8) :: pdelt, pt0(NLEV), pdsigma(NLEV), pg(NLEV, NLEV), ptau(NLEV, NLEV)
  real(8) :: pbm1(NLEV, NLEV, NTRU)
  delt = pdelt
  t0(:) = pt0(:)
  dsigma(:) = pdsigma(:)
  g(:, :) = pg(:, :)
  tau(:, :) = ptau(:, :)
  call makebm
  pbm1(:, :, :) = bm1(:, :, :)
end subroutine

! sections 1-3 of spectrala (src/plasim.f90:2817-2876) through the routine
! itself: the state at t-dt (spm, sdm, stm), the tendencies (sdt, stt, spt),
! the orography and the matrices in; sdt, stt, spt out.  The rest of
! spectrala (time filter, tendencies added into the *p arrays) does not
! write these three.  Spectral arrays are (NSPP, NLEV) column-major.
subroutine ref_plasim_si(pdelt, pt0, pdsigma, pg, ptau, pbm1, psop, pspm, patm, padm, &
                         pstt, psdt, pspt) bind(C, name="ref_plasim_si")
  use pumamod
  real(8) :: pdelt, pt0(NLEV), pdsigma(NLEV), pg(NLEV, NLEV), ptau(NLEV, NLEV)
  real(8) :: pbm1(NLEV, NLEV, NTRU)
  real(8) :: psop(NSPP), pspm(NSPP), patm(NSPP, NLEV), padm(NSPP, NLEV)
  real(8) :: pstt(NSPP, NLEV), psdt(NSPP, NLEV), pspt(NSPP)
  integer :: jm, jn, jr
  delt = pdelt
  delt2 = 2.0 * pdelt
  t0(:) = pt0(:)
  dsigma(:) = pdsigma(:)
  g(:, :) = pg(:, :)
  tau(:, :) = ptau(:, :)
  bm1(:, :, :) = pbm1(:, :, :)
  jr = -1
  do jm = 0, NTRU
    do jn = jm, NTRU
      jr = jr + 2
      nindex(jr) = jn
      nindex(jr + 1) = jn
    enddo
  enddo
  sop(:) = psop(:)
  spm(:) = pspm(:)
  stm(:, :) = patm(:, :)
  sdm(:, :) = padm(:, :)
  szm(:, :) = 0.0
  sqm(:, :) = 0.0
  stt(:, :) = pstt(:, :)
  sdt(:, :) = psdt(:, :)
  spt(:) = pspt(:)
  szt(:, :) = 0.0
  sqt(:, :) = 0.0
  nadv = 1
  nkits = 1
  mrnum = 0
  ndiagsp = 0
  nenergy = 0
  nentropy = 0
  call spectrala
  pstt(:, :) = stt(:, :)
  psdt(:, :) = sdt(:, :)
  pspt(:) = spt(:)
end subroutine

! calcgp (src/plasim.f90:2587-2750): the grid-point fields of gridpointa
! (gu, gv, gt, gd, gz, gq, gp = ps/psurf, gpj, gpm = the Fourier-space
! d/dlambda of ln ps, dq, rcsq) and the vertical set-up of ref_plasim_initsi
! in; gtn, gqn, guz, gvz, gvp, gphi and dw out.  Grid arrays are
! (NHOR, NLEV) column-major.  scal = (rdbrv, adv, psurf, ww); kq = nqspec.
subroutine ref_plasim_calcgp(pgu, pgv, pgt, pgd, pgz, pgq, pgp, pgpj, pgpm, pdq, prcsq, &
                             psigmah, pdsigma, prdsig, pt0, pakap, pscal, kq, &
                             pgtn, pgqn, pguz, pgvz, pgvp, pgphi, pdw) &
    bind(C, name="ref_plasim_calcgp")
  use pumamod
  real(8) :: pgu(NHOR, NLEV), pgv(NHOR, NLEV), pgt(NHOR, NLEV), pgd(NHOR, NLEV)
  real(8) :: pgz(NHOR, NLEV), pgq(NHOR, NLEV), pgp(NHOR), pgpj(NHOR), pgpm(NHOR)
  real(8) :: pdq(NHOR, NLEP), prcsq(NHOR)
  real(8) :: psigmah(NLEV), pdsigma(NLEV), prdsig(NLEV), pt0(NLEV), pakap, pscal(4)
  integer(4), value :: kq
  real(8) :: pgtn(NHOR, NLEV), pgqn(NHOR, NLEV), pguz(NHOR, NLEV), pgvz(NHOR, NLEV)
  real(8) :: pgvp(NHOR), pgphi(NHOR, NLEV), pdw(NHOR, NLEV)
  gu(:, :) = pgu(:, :)
  gv(:, :) = pgv(:, :)
  gt(:, :) = pgt(:, :)
  gd(:, :) = pgd(:, :)
  gz(:, :) = pgz(:, :)
  gq(:, :) = pgq(:, :)
  gp(:) = pgp(:)
  gpj(:) = pgpj(:)
  dq(:, :) = pdq(:, :)
  rcsq(:) = prcsq(:)
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
  call calcgp(pgtn, pgqn, pguz, pgvz, pgpm, pgvp, pgphi)
  pdw(:, :) = dw(:, :)
end subroutine
