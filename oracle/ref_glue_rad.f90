! ref_glue_rad.f90 -- TEST INFRASTRUCTURE ONLY (part of the oracle).
!
! Calls the reference's long-wave column (radmod lwr, compiled from its
! sources at the default resolution module: NHOR = 2048 columns, NLEV = 10)
! on columns given from C, with the pumamod/radmod inputs set as the model
! sets them before radstep calls lwr.  Arrays are (NHOR, lev) column-major.

subroutine ref_rad_dims(d) bind(C, name="ref_rad_dims")
  use radmod
  integer(4) :: d(2)
  d(1) = NHOR
  d(2) = NLEV
end subroutine

subroutine ref_rad_lwr(xsig, xsigh, xdsig, pga, pth2oc, ptpofmt, pacllwr, pclgray, npb, &
                       pdp, pdt, pdq, pqo3, pqco2, pcc, pql, pls, pftu, pftd) &
                       bind(C, name="ref_rad_lwr")
  use radmod
  real(8) :: xsig(NLEV), xsigh(NLEV), xdsig(NLEV), pga, pth2oc, ptpofmt, pacllwr, pclgray
  integer(4), value :: npb
  real(8) :: pdp(NHOR), pdt(NHOR, NLEP), pdq(NHOR, NLEV), pqo3(NHOR, NLEV)
  real(8) :: pqco2(NHOR, NLEV), pcc(NHOR, NLEV), pql(NHOR, NLEV), pls(NHOR)
  real(8) :: pftu(NHOR, NLEP), pftd(NHOR, NLEP)
  sigma(:) = xsig(:)
  sigmah(:) = xsigh(:)
  dsigma(:) = xdsig(:)
  ga = pga
  th2oc = pth2oc
  tpofmt = ptpofmt
  acllwr = pacllwr
  clgray = pclgray
  npbroaden = npb
  dp(:) = pdp(:)
  dt(:, :) = pdt(:, :)
  dq(:, 1:NLEV) = pdq(:, :)
  dqo3(:, :) = pqo3(:, :)
  dqco2(:, :) = pqco2(:, :)
  dcc(:, 1:NLEV) = pcc(:, :)
  dql(:, 1:NLEV) = pql(:, :)
  dls(:) = pls(:)
  call lwr
  pftu(:, :) = dftu(:, :)
  pftd(:, :) = dftd(:, :)
end subroutine
