! exo_transforms.f90 -- drop-in replacements of the reference's transform
! subroutines (exoplasim/plasim/src/legmod.f90, fftmod.f90, gaussmod.f90)
! with the same names and dummy arguments, computed on the GPU by
! libexosht.so.  The arrays keep the reference layout; `real` is the model's
! kind (MOST_PREC), selected at compile time.  Sizes come from the tables
! built by legini (exo_legini_model.f90 or exo_legini called directly).

! inigau (src/gaussmod.f90:5)
subroutine inigau(klat, pz0, pzw)
  use exo_legmod_capi
  implicit none
  integer :: klat
  real(kind=8), target :: pz0(klat), pzw(klat)
  call exo_check(exo_inigau(int(klat, c_int), c_loc(pz0(1)), c_loc(pzw(1))), 'inigau')
end subroutine inigau

! sp2fc (src/legmod.f90:230)
subroutine sp2fc(sp, fc)
  use exo_legmod_capi
  implicit none
  real, target :: sp(*), fc(*)
  if (kind(sp) == 8) then
    call exo_check(exo_sp2fc_f64(c_loc(sp(1)), c_loc(fc(1)), 1_c_int), 'sp2fc')
  else
    call exo_check(exo_sp2fc_f32(c_loc(sp(1)), c_loc(fc(1)), 1_c_int), 'sp2fc')
  endif
end subroutine sp2fc

! sp2fcdmu (src/legmod.f90:262)
subroutine sp2fcdmu(sp, fc)
  use exo_legmod_capi
  implicit none
  real, target :: sp(*), fc(*)
  if (kind(sp) == 8) then
    call exo_check(exo_sp2fcdmu_f64(c_loc(sp(1)), c_loc(fc(1)), 1_c_int), 'sp2fcdmu')
  else
    call exo_check(exo_sp2fcdmu_f32(c_loc(sp(1)), c_loc(fc(1)), 1_c_int), 'sp2fcdmu')
  endif
end subroutine sp2fcdmu

! sp2fl (src/legmod.f90:613): sp2fc of klev consecutive levels
subroutine sp2fl(psp, pfc, klev)
  use exo_legmod_capi
  implicit none
  integer :: klev
  real, target :: psp(*), pfc(*)
  if (kind(psp) == 8) then
    call exo_check(exo_sp2fc_f64(c_loc(psp(1)), c_loc(pfc(1)), int(klev, c_int)), 'sp2fl')
  else
    call exo_check(exo_sp2fc_f32(c_loc(psp(1)), c_loc(pfc(1)), int(klev, c_int)), 'sp2fl')
  endif
end subroutine sp2fl

! fc2sp (src/legmod.f90:178)
subroutine fc2sp(fc, sp)
  use exo_legmod_capi
  implicit none
  real, target :: fc(*), sp(*)
  if (kind(sp) == 8) then
    call exo_check(exo_fc2sp_f64(c_loc(fc(1)), c_loc(sp(1)), 1_c_int), 'fc2sp')
  else
    call exo_check(exo_fc2sp_f32(c_loc(fc(1)), c_loc(sp(1)), 1_c_int), 'fc2sp')
  endif
end subroutine fc2sp

! dv2uv (src/legmod.f90:312); pz is not modified
subroutine dv2uv(pd, pz, pu, pv)
  use exo_legmod_capi
  implicit none
  real, target :: pd(*), pz(*), pu(*), pv(*)
  if (kind(pd) == 8) then
    call exo_check(exo_dv2uv_f64(c_loc(pd(1)), c_loc(pz(1)), c_loc(pu(1)), c_loc(pv(1))), 'dv2uv')
  else
    call exo_check(exo_dv2uv_f32(c_loc(pd(1)), c_loc(pz(1)), c_loc(pu(1)), c_loc(pv(1))), 'dv2uv')
  endif
end subroutine dv2uv

! uv2dv (src/legmod.f90:356)
subroutine uv2dv(pu, pv, pd, pz)
  use exo_legmod_capi
  implicit none
  real, target :: pu(*), pv(*), pd(*), pz(*)
  if (kind(pd) == 8) then
    call exo_check(exo_uv2dv_f64(c_loc(pu(1)), c_loc(pv(1)), c_loc(pd(1)), c_loc(pz(1))), 'uv2dv')
  else
    call exo_check(exo_uv2dv_f32(c_loc(pu(1)), c_loc(pv(1)), c_loc(pd(1)), c_loc(pz(1))), 'uv2dv')
  endif
end subroutine uv2dv

! qtend (src/legmod.f90:433)
subroutine qtend(q, qn, uq, vq)
  use exo_legmod_capi
  implicit none
  real, target :: q(*), qn(*), uq(*), vq(*)
  if (kind(q) == 8) then
    call exo_check(exo_qtend_f64(c_loc(q(1)), c_loc(qn(1)), c_loc(uq(1)), c_loc(vq(1))), 'qtend')
  else
    call exo_check(exo_qtend_f32(c_loc(q(1)), c_loc(qn(1)), c_loc(uq(1)), c_loc(vq(1))), 'qtend')
  endif
end subroutine qtend

! mktend (src/legmod.f90:505)
subroutine mktend(d, t, z, tn, fu, fv, ke, ut, vt)
  use exo_legmod_capi
  implicit none
  real, target :: d(*), t(*), z(*), tn(*), fu(*), fv(*), ke(*), ut(*), vt(*)
  if (kind(d) == 8) then
    call exo_check(exo_mktend_f64(c_loc(d(1)), c_loc(t(1)), c_loc(z(1)), c_loc(tn(1)), &
                   c_loc(fu(1)), c_loc(fv(1)), c_loc(ke(1)), c_loc(ut(1)), c_loc(vt(1))), 'mktend')
  else
    call exo_check(exo_mktend_f32(c_loc(d(1)), c_loc(t(1)), c_loc(z(1)), c_loc(tn(1)), &
                   c_loc(fu(1)), c_loc(fv(1)), c_loc(ke(1)), c_loc(ut(1)), c_loc(vt(1))), 'mktend')
  endif
end subroutine mktend

! gp2fc (src/fftmod.f90:29, fft991mod.f90:32), in place
subroutine gp2fc(a, n, lot)
  use exo_legmod_capi
  implicit none
  integer :: n, lot
  real, target :: a(*)
  if (kind(a) == 8) then
    call exo_check(exo_gp2fc_f64(c_loc(a(1)), int(n, c_int), int(lot, c_int)), 'gp2fc')
  else
    call exo_check(exo_gp2fc_f32(c_loc(a(1)), int(n, c_int), int(lot, c_int)), 'gp2fc')
  endif
end subroutine gp2fc

! fc2gp (src/fftmod.f90:64, fft991mod.f90:60), in place
subroutine fc2gp(a, n, lot)
  use exo_legmod_capi
  implicit none
  integer :: n, lot
  real, target :: a(*)
  if (kind(a) == 8) then
    call exo_check(exo_fc2gp_f64(c_loc(a(1)), int(n, c_int), int(lot, c_int)), 'fc2gp')
  else
    call exo_check(exo_fc2gp_f32(c_loc(a(1)), int(n, c_int), int(lot, c_int)), 'fc2gp')
  endif
end subroutine fc2gp

! fftini (src/fftmod.f90:93-122, fft991mod.f90:87-112): the check of the
! FFT length.  The twiddles live in the GPU plan (built by legini), so only
! the check remains: an unsupported n ends the run through the abort path,
! as the reference's stop does.
subroutine fftini(n)
  use exo_legmod_capi
  implicit none
  integer :: n
  character(len=80) :: msg
  if (n /= 64 .and. n /= 128 .and. n /= 192 .and. n /= 384) then
    write(msg, '(a,i0,a)') 'exoplasim_amd: FFT does not support n = ', n, &
        ' (NLON 64, 128, 192, 384)'
    call exo_abort(msg)
  endif
end subroutine fftini

! ql (src/gaussmod.f90:46-69): the Legendre polynomial of degree k at p,
! normalised to a unit integral of its square over [-1, 1], by the upward
! three-term recurrence of the normalised polynomials
real(kind=8) function ql(k, p)
  implicit none
  integer :: k
  real(kind=8) :: p
  real(kind=8) :: z0, z1, z2
  integer :: j
  z0 = sqrt(0.5_8)
  if (k == 0) then
    ql = z0
    return
  endif
  z1 = sqrt(1.5_8) * p
  do j = 2, k
    z2 = sqrt((2.0_8 * j + 1.0_8) * (2.0_8 * j - 1.0_8)) / j * p * z1 &
         - (j - 1.0_8) / j * sqrt((2.0_8 * j + 1.0_8) / (2.0_8 * j - 3.0_8)) * z0
    z0 = z1
    z1 = z2
  enddo
  ql = z1
end function ql

! qld (src/gaussmod.f90:75-86): the reciprocal of dql/dp, the Newton factor
! of inigau, from (p^2 - 1) dP_k/dp = k (p P_k - c_k P_{k-1}) for the
! normalised polynomials, c_k = sqrt((2k+1)/(2k-1))
real(kind=8) function qld(k, p)
  implicit none
  integer :: k
  real(kind=8) :: p
  real(kind=8), external :: ql
  qld = (p * p - 1.0_8) / (k * (p * ql(k, p) - sqrt((2.0_8 * k + 1.0_8) / &
        (2.0_8 * k - 1.0_8)) * ql(k - 1, p)))
end function qld
