// dyn_calcgp.hip -- the grid-point dynamics of PlaSim between the inverse and
// the direct transforms, on the GPU (declared in include/exo_dyn.h):
//
//   k_calcgp  calcgp (src/plasim.f90:2587-2750): virtual temperatures,
//             vertical velocity (sigma-dot), the nonlinear tendencies gtn,
//             gqn, guz, gvz, the geopotential term gphi and dw; with it the
//             products of gridpointa (src/plasim.f90:2436-2446) and, on
//             request, gp = exp(gp) (:2434)
//   k_gpmt    the Fourier-space d/dlambda of ln(ps) (src/plasim.f90:2425-2430)
//
// One thread per grid point and member runs down the column.  calcgp is a
// pointwise column computation: every level of every field is read once and
// written once, coalesced along the grid points, so the kernel is bound by
// HBM; the two triangular sums over levels (ztpta/ztptb against c, gphi
// against g) run from registers, with c and g read through scalar loads
// (every lane of a wave reads the same element).
#include "../../include/exo_dyn.h"
#include "sht_plan.h"

namespace exo {

template <typename R>
struct DynVert {
  const R *sigmah, *dsigma, *rdsig, *t0, *t01s2, *tkp, *c, *g;
  R akap, rdbrv, adv, psurf, ww;
};

template <typename R>
struct CalcgpPtrs {
  const R *gu, *gv, *gt, *gd, *gz, *gq, *dq;
  R* gp;
  const R *gpj, *gpm, *rcsq;
  R *gtn, *gqn, *guz, *gvz, *gphi, *dw, *gvp;
  R *gut, *gvt, *gke, *guq, *gvq;
};

// NL > 0: exactly NL levels, every per-level array in registers; NL = 0:
// a.nlev <= 64 levels at run time, the arrays in per-lane memory.
template <typename R, int NL>
__global__ __launch_bounds__(256) void k_calcgp(CalcgpPtrs<R> f, DynVert<R> v, int nlev_rt,
                                                int nhor, long mstride, int nqspec, int exp_gp) {
  constexpr int NA = NL > 0 ? NL : 64;
  const int n = NL > 0 ? NL : nlev_rt;
  const int h = blockIdx.x * 256 + threadIdx.x;
  if (h >= nhor) return;
  const long mo = (long)blockIdx.y * mstride;
  auto at = [&](const R* p, int j) -> R { return p[mo + (long)j * nhor + h]; };
  auto put = [&](R* p, int j, R x) { p[mo + (long)j * nhor + h] = x; };
  const R gpm = f.gpm[mo + h], gpj = f.gpj[mo + h], rcsq = f.rcsq[mo + h];
  R gp = f.gp[mo + h];
  if (exp_gp) {
    gp = exp(gp);
    f.gp[mo + h] = gp;
  }
  const R* __restrict__ dsig = v.dsigma;
  const R* __restrict__ sigh = v.sigmah;
  const R* __restrict__ c = v.c;
  const R* __restrict__ g = v.g;

  // zvgpg = rcsq (u d/dlambda + v d/dmu) ln ps; the running sums of the
  // sigma-dot integral in the reference's order
  R zvgpg[NA], gdl[NA], ztv1d[NA], ke[NA];
  R zsumd = R(0), vp = R(0);
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    if (j >= n) break;
    const R gu = at(f.gu, j), gv = at(f.gv, j), gd = at(f.gd, j);
    zvgpg[j] = rcsq * (gu * gpm + gv * gpj);
    gdl[j] = gd;
    zsumd = zsumd + dsig[j] * gd;
    vp = vp + dsig[j] * zvgpg[j];
  }
  f.gvp[mo + h] = vp;
  const R tot = zsumd + vp;  // zsumd + gvp at the bottom
  const R qfac = R(1) / v.rdbrv - R(1);

  // level j: virtual temperatures, then the terms of level j that need
  // level j + 1 (gtd, gqm, gud, gvd), carried one level on
  R rz = R(0), rv = R(0);  // running sums again, for zsdotp(j)
  R gtd_m = R(0), gqm_m = R(0), gud_m = R(0), gvd_m = R(0);
  R zs2 = R(0);            // the second zsumd of the reference (:2656, :2681)
  R gt = at(f.gt, 0), gq = at(f.gq, 0), gu = at(f.gu, 0), gv = at(f.gv, 0);
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    if (j >= n) break;
    const R gz = at(f.gz, j);
    R zq = (nqspec == 1) ? gq / gp : at(f.dq, j);
    if (!(zq > R(0))) zq = R(0);
    // The reference forms (gt + t0)(1 + (1/rdbrv - 1) zq) - t0 and then
    // subtracts gt again for gphi (src/plasim.f90:2617-2622, 2742): t0 is
    // large against the humidity term, and the subtractions cancel it.  The
    // same quantities in the form without the cancellation: the virtual
    // temperature term corr = (gt + t0)(1/rdbrv - 1) zq, ztv1 = gt + corr,
    // ztv1 - gt = corr and ztv2 = (gt + corr - t0 adv zq) / (1 + adv zq).
    const R t0 = v.t0[j];
    const R corr = (gt + t0) * (qfac * zq);
    const R ztv1 = gt + corr;
    const R ztv2 = (gt + corr - t0 * (v.adv * zq)) / (R(1) + v.adv * zq);
    ztv1d[j] = corr;
    // sigma-dot and the vertical advection terms at the interface below j
    R gtd = R(0), gqm = R(0), gud = R(0), gvd = R(0);
    R gt1 = R(0), gq1 = R(0), gu1 = R(0), gv1 = R(0);
    if (j < n - 1) {
      rz = rz + dsig[j] * gdl[j];
      rv = rv + dsig[j] * zvgpg[j];
      const R zsdotp = sigh[j] * tot - (rz + rv);
      gt1 = at(f.gt, j + 1);
      gq1 = at(f.gq, j + 1);
      gu1 = at(f.gu, j + 1);
      gv1 = at(f.gv, j + 1);
      gtd = zsdotp * (gt1 - gt);
      gqm = zsdotp * (gq1 + gq);
      gud = zsdotp * (gu1 - gu);
      gvd = zsdotp * (gv1 - gv);
    }
    const R rd = v.rdsig[j];
    R gtn, gqn, guz, gvz, dw;
    if (j == 0) {
      zs2 = zvgpg[0] * dsig[0];
      gtn = gt * gdl[0] - v.akap * ztv2 * gdl[0] -
            rd * (gtd + v.t01s2[0] * (sigh[0] * vp - zs2));
      gqn = -rd * gqm;
      guz = -gu * gz - gpj * ztv1 - rd * gvd;
      gvz = gv * gz - gpm * ztv1 - rd * gud;
      dw = gdl[0] * gp * v.psurf * v.ww;
    } else {
      // c(k, j) for k <= j: column j of c, through scalar loads
      R ztpta = c[j * n] * zvgpg[0];
      R ztptb = c[j * n] * (zvgpg[0] + gdl[0]);
#pragma unroll
      for (int k = 1; k < NA; ++k) {
        if (k > j) break;
        ztpta = ztpta + c[j * n + k] * zvgpg[k];
        ztptb = ztptb + c[j * n + k] * (zvgpg[k] + gdl[k]);
      }
      if (j < n - 1) {
        zs2 = zs2 + zvgpg[j] * dsig[j];
        gtn = gt * gdl[j] + v.akap * ztv2 * (zvgpg[j] - ztptb) + v.tkp[j] * (zvgpg[j] - ztpta) -
              rd * (gtd + gtd_m + vp * (v.t01s2[j] * sigh[j] + v.t01s2[j - 1] * sigh[j - 1]) -
                    zs2 * (v.t01s2[j - 1] + v.t01s2[j]) + zvgpg[j] * dsig[j] * v.t01s2[j - 1]);
        gqn = -rd * (gqm - gqm_m);
        guz = -gu * gz - gpj * ztv1 - rd * (gvd + gvd_m);
        gvz = gv * gz - gpm * ztv1 - rd * (gud + gud_m);
      } else {
        gtn = gt * gdl[j] + v.akap * ztv2 * (zvgpg[j] - ztptb) + v.tkp[j] * (zvgpg[j] - ztpta) -
              rd * (gtd_m + v.t01s2[j - 1] * (sigh[j - 1] * vp - zs2));
        gqn = rd * gqm_m;
        guz = -gu * gz - gpj * ztv1 - rd * gvd_m;
        gvz = gv * gz - gpm * ztv1 - rd * gud_m;
      }
      dw = (zvgpg[j] - ztptb) * gp * v.psurf * v.ww;
    }
    put(f.gtn, j, gtn);
    put(f.gqn, j, gqn);
    put(f.guz, j, guz);
    put(f.gvz, j, gvz);
    put(f.dw, j, dw);
    if (f.gut) put(f.gut, j, gu * gt);
    if (f.gvt) put(f.gvt, j, gv * gt);
    if (nqspec == 1 && f.guq) put(f.guq, j, gu * gq);
    if (nqspec == 1 && f.gvq) put(f.gvq, j, gv * gq);
    ke[j] = gu * gu + gv * gv;
    gtd_m = gtd;
    gqm_m = gqm;
    gud_m = gud;
    gvd_m = gvd;
    gt = gt1;
    gq = gq1;
    gu = gu1;
    gv = gv1;
  }
  // gphi(j) = 2 / rcsq * sum_k g(k, j) (ztv1 - gt)(k)
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    if (j >= n) break;
    R s = R(0);
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      if (k >= n) break;
      s = s + g[j * n + k] * ztv1d[k];
    }
    const R gphi = s * R(2) / rcsq;
    put(f.gphi, j, gphi);
    if (f.gke) put(f.gke, j, ke[j] + gphi);
  }
}

// rcsq = 1 / cos^2 per grid point from the Gaussian abscissas, as the model
// sets it up (src/plasim.f90:1897 csq = 1 - sid^2 in the real kind, :205)
template <typename R>
__global__ void k_rcsq(const double* sid, int nlat, int nlon, R* out, long mstride) {
  const int h = blockIdx.x * 256 + threadIdx.x;
  if (h >= nlat * nlon) return;
  const R csq = (R)(1.0 - sid[h / nlon] * sid[h / nlon]);
  out[blockIdx.y * mstride + h] = R(1) / csq;
}

// Fourier-space d/dlambda: one complex coefficient (m) per thread
template <typename R>
__global__ void k_gpmt(const R* in, R* out, int nlon, long total) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int m = (int)(i % (nlon / 2));
  const R re = in[2 * i], im = in[2 * i + 1];
  out[2 * i] = -im * (R)m;
  out[2 * i + 1] = re * (R)m;
}

}  // namespace exo

using namespace exo;

struct exo_dyn_s {
  int nlev = 0;
  double* d64 = nullptr;  // sigmah dsigma rdsig t0 t01s2 tkp (nlev each), c, g (nlev^2 each)
  float* d32 = nullptr;
  double sc[5] = {0, 0, 0, 0, 0};  // akap rdbrv adv psurf ww
};

template <typename R>
static DynVert<R> vert_of(const exo_dyn_s* d) {
  const int n = d->nlev;
  const R* b;
  if constexpr (sizeof(R) == 8)
    b = d->d64;
  else
    b = d->d32;
  DynVert<R> v;
  v.sigmah = b;
  v.dsigma = b + n;
  v.rdsig = b + 2 * n;
  v.t0 = b + 3 * n;
  v.t01s2 = b + 4 * n;
  v.tkp = b + 5 * n;
  v.c = b + 6 * n;
  v.g = b + 6 * n + n * n;
  v.akap = (R)d->sc[0];
  v.rdbrv = (R)d->sc[1];
  v.adv = (R)d->sc[2];
  v.psurf = (R)d->sc[3];
  v.ww = (R)d->sc[4];
  return v;
}

extern "C" int exo_dyn_create(exo_dyn_t* dyn, int nlev, const double* sigmah,
                              const double* dsigma, const double* rdsig, const double* t0,
                              const double* t01s2, const double* tkp, const double* c,
                              const double* g, double akap, double rdbrv, double adv,
                              double psurf, double ww) {
  if (!dyn) return EXO_ERR_ARG;
  *dyn = nullptr;
  if (nlev < 2 || nlev > 64 || !sigmah || !dsigma || !rdsig || !t0 || !t01s2 || !tkp || !c ||
      !g)
    return EXO_ERR_ARG;
  exo_dyn_s* d = new exo_dyn_s();
  d->nlev = nlev;
  d->sc[0] = akap;
  d->sc[1] = rdbrv;
  d->sc[2] = adv;
  d->sc[3] = psurf;
  d->sc[4] = ww;
  const int tot = 6 * nlev + 2 * nlev * nlev;
  double* host = new double[tot];
  const double* parts[6] = {sigmah, dsigma, rdsig, t0, t01s2, tkp};
  for (int k = 0; k < 6; ++k)
    for (int i = 0; i < nlev; ++i) host[k * nlev + i] = parts[k][i];
  for (int i = 0; i < nlev * nlev; ++i) {
    host[6 * nlev + i] = c[i];
    host[6 * nlev + nlev * nlev + i] = g[i];
  }
  float* host32 = new float[tot];
  for (int i = 0; i < tot; ++i) host32[i] = (float)host[i];
  int rc = EXO_OK;
  if (hipMalloc(&d->d64, sizeof(double) * tot) != hipSuccess ||
      hipMalloc(&d->d32, sizeof(float) * tot) != hipSuccess ||
      hipMemcpy(d->d64, host, sizeof(double) * tot, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d->d32, host32, sizeof(float) * tot, hipMemcpyHostToDevice) != hipSuccess)
    rc = EXO_ERR_HIP;
  delete[] host;
  delete[] host32;
  if (rc != EXO_OK) {
    exo_dyn_destroy(d);
    return rc;
  }
  *dyn = d;
  return EXO_OK;
}

extern "C" int exo_dyn_destroy(exo_dyn_t d) {
  if (!d) return EXO_ERR_ARG;
  (void)hipDeviceSynchronize();
  if (d->d64) (void)hipFree(d->d64);
  if (d->d32) (void)hipFree(d->d32);
  delete d;
  return EXO_OK;
}

template <typename R>
static int calcgp_t(exo_dyn_t d, int nmem, long ms, int nhor, int nqspec, int exp_gp,
                    const exo_calcgp_fields* f, hipStream_t s) {
  CalcgpPtrs<R> p;
  p.gu = (const R*)f->gu;
  p.gv = (const R*)f->gv;
  p.gt = (const R*)f->gt;
  p.gd = (const R*)f->gd;
  p.gz = (const R*)f->gz;
  p.gq = (const R*)f->gq;
  p.dq = (const R*)f->dq;
  p.gp = (R*)f->gp;
  p.gpj = (const R*)f->gpj;
  p.gpm = (const R*)f->gpm;
  p.rcsq = (const R*)f->rcsq;
  p.gtn = (R*)f->gtn;
  p.gqn = (R*)f->gqn;
  p.guz = (R*)f->guz;
  p.gvz = (R*)f->gvz;
  p.gphi = (R*)f->gphi;
  p.dw = (R*)f->dw;
  p.gvp = (R*)f->gvp;
  p.gut = (R*)f->gut;
  p.gvt = (R*)f->gvt;
  p.gke = (R*)f->gke;
  p.guq = (R*)f->guq;
  p.gvq = (R*)f->gvq;
  const DynVert<R> v = vert_of<R>(d);
  dim3 grid((unsigned)((nhor + 255) / 256), (unsigned)nmem);
  switch (d->nlev) {
    case 10: k_calcgp<R, 10><<<grid, 256, 0, s>>>(p, v, d->nlev, nhor, ms, nqspec, exp_gp); break;
    case 20: k_calcgp<R, 20><<<grid, 256, 0, s>>>(p, v, d->nlev, nhor, ms, nqspec, exp_gp); break;
    case 40: k_calcgp<R, 40><<<grid, 256, 0, s>>>(p, v, d->nlev, nhor, ms, nqspec, exp_gp); break;
    default: k_calcgp<R, 0><<<grid, 256, 0, s>>>(p, v, d->nlev, nhor, ms, nqspec, exp_gp); break;
  }
  return hipGetLastError() == hipSuccess ? EXO_OK : EXO_ERR_LAUNCH;
}

extern "C" int exo_dyn_calcgp(exo_dyn_t d, int mode, int nmem, long member_stride, int nhor,
                              int nqspec, int exp_gp, const exo_calcgp_fields* f,
                              void* stream) {
  if (!d || !f || mode < EXO_F64 || mode > EXO_BF16X3 || nmem < 0 || nhor < 0 ||
      (nqspec != 0 && nqspec != 1))
    return EXO_ERR_ARG;
  if (nmem == 0 || nhor == 0) return EXO_OK;
  if (nmem > 1 && member_stride <= 0) return EXO_ERR_ARG;
  const void* need[] = {f->gu, f->gv, f->gt, f->gd, f->gz, f->gq, f->gp, f->gpj, f->gpm,
                        f->rcsq, f->gtn, f->gqn, f->guz, f->gvz, f->gphi, f->dw, f->gvp};
  for (const void* q : need)
    if (!q) return EXO_ERR_ARG;
  if (nqspec == 0 && !f->dq) return EXO_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (mode == EXO_F64) return calcgp_t<double>(d, nmem, member_stride, nhor, nqspec, exp_gp, f, s);
  return calcgp_t<float>(d, nmem, member_stride, nhor, nqspec, exp_gp, f, s);
}

extern "C" int exo_dyn_gpmt(int mode, int nlon, long nrow, const void* in, void* out,
                            void* stream) {
  if (mode < EXO_F64 || mode > EXO_BF16X3 || nlon < 2 || (nlon & 1) || nrow < 0 ||
      (nrow > 0 && (!in || !out)))
    return EXO_ERR_ARG;
  const long total = nrow * (nlon / 2);
  if (total == 0) return EXO_OK;
  const unsigned nb = (unsigned)((total + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  if (mode == EXO_F64)
    k_gpmt<double><<<nb, 256, 0, s>>>((const double*)in, (double*)out, nlon, total);
  else
    k_gpmt<float><<<nb, 256, 0, s>>>((const float*)in, (float*)out, nlon, total);
  return hipGetLastError() == hipSuccess ? EXO_OK : EXO_ERR_LAUNCH;
}

// ---------------------------------------------------------------------------
// the device-resident gridpointa transform group
// ---------------------------------------------------------------------------
// work slots (units of NHOR values) of one member
namespace {
struct GpaSlots {
  int gz, gd, gt, gq, gp, gpj, gpm, gu, gv, dq, rcsq;
  int a;  // first FFT slot: gtn gqn guz gvz gut gvt gke guq gvq (nlev each), gvp
  int gtn, gqn, guz, gvz, gut, gvt, gke, guq, gvq, gvp, gphi, dw, total;
  explicit GpaSlots(int n) {
    gz = 0, gd = n, gt = 2 * n, gq = 3 * n, gp = 4 * n, gpj = 4 * n + 1, gpm = 4 * n + 2;
    gu = 4 * n + 3, gv = 5 * n + 3, dq = 6 * n + 3, rcsq = 7 * n + 3;
    a = 7 * n + 4;
    gtn = a, gqn = a + n, guz = a + 2 * n, gvz = a + 3 * n, gut = a + 4 * n, gvt = a + 5 * n;
    gke = a + 6 * n, guq = a + 7 * n, gvq = a + 8 * n, gvp = a + 9 * n;
    gphi = a + 9 * n + 1, dw = a + 10 * n + 1, total = a + 11 * n + 1;
  }
};
}  // namespace

extern "C" int exo_gridpointa_fields(int nlev) {
  return nlev < 2 ? EXO_ERR_ARG : GpaSlots(nlev).total;
}

extern "C" int exo_gridpointa(exo_sht_plan_t p, exo_dyn_t d, int mode, int nmem, int nqspec,
                              const void* sp_state, long sp_stride, void* work, long work_stride,
                              void* sp_tend, long tend_stride, void* stream) {
  if (!p || !d || mode < EXO_F64 || mode > EXO_BF16X3 || nmem < 0 ||
      (nqspec != 0 && nqspec != 1))
    return EXO_ERR_ARG;
  if (nmem == 0) return EXO_OK;
  const Tables& tb = p->tb;
  const int n = d->nlev, nsp = 4 * n + 1;
  const long nhor = (long)tb.nlat * tb.nlon;
  const GpaSlots sl(n);
  if (!sp_state || !work || !sp_tend || sp_stride < (long)nsp * 2 * tb.ncsp ||
      tend_stride < (long)nsp * 2 * tb.ncsp || work_stride < (long)sl.total * nhor)
    return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(p->mu);
  hipStream_t s = (hipStream_t)stream;
  const size_t es = (mode == MODE_F64) ? 8 : 4;
  char* wk = (char*)work;
  auto slot = [&](int k) { return (void*)(wk + es * nhor * k); };

  // 1 / cos^2 per grid point
  {
    const dim3 g((unsigned)((nhor + 255) / 256), (unsigned)nmem);
    if (mode == MODE_F64)
      k_rcsq<double><<<g, 256, 0, s>>>(tb.sid, tb.nlat, tb.nlon, (double*)slot(sl.rcsq),
                                        work_stride);
    else
      k_rcsq<float><<<g, 256, 0, s>>>(tb.sid, tb.nlat, tb.nlon, (float*)slot(sl.rcsq),
                                       work_stride);
    if (hipGetLastError() != hipSuccess) return EXO_ERR_LAUNCH;
  }

  // invlega (src/legmod.f90:635-656) fused with fc2gp, and gpmt
  InvArgs base{};
  base.sp = sp_state;
  base.out = work;
  base.sp_stride = sp_stride;
  base.out_stride = work_stride;
  base.nmem = nmem;
  base.plavor = tb.plavor;
  base.var = p->inv_var[mode];
  if (mode == MODE_BF16X3) {
    const int rc = plan_grow(&p->pk, &p->pk_bytes, pack_bytes(tb, nmem, nsp));
    if (rc) return rc;
    if (launch_pack(tb, (const float*)sp_state, sp_stride, 0, nsp, nmem, (float*)p->pk, s))
      return EXO_ERR_LAUNCH;
    base.bpk = p->pk;
    base.bpk_stride = (long)nsp * pk_field(tb.ntru);
    base.bpk_f0 = 0;
  }
  auto scalar = [&](int table, int src0, int dst0, int nfield, int dlam) {
    InvArgs a = base;
    a.kind = INV_SCALAR;
    a.table = table;
    a.src0 = src0;
    a.dst0 = dst0;
    a.nfield = nfield;
    a.dlam = dlam;
    return launch_inverse(tb, mode, true, a, s);
  };
  // sz sd st (sq, sp) -> gz gd gt (gq, gp); with nqspec = 0 gq is the caller's
  if (scalar(TQI, 0, sl.gz, nqspec ? nsp : 3 * n, 0)) return EXO_ERR_LAUNCH;
  if (!nqspec && scalar(TQI, 4 * n, sl.gp, 1, 0)) return EXO_ERR_LAUNCH;
  if (scalar(TQJ, 4 * n, sl.gpj, 1, 0)) return EXO_ERR_LAUNCH;
  if (scalar(TQI, 4 * n, sl.gpm, 1, 1)) return EXO_ERR_LAUNCH;
  {
    InvArgs a = base;
    a.kind = INV_UV;
    a.var = p->uv_var[mode];
    a.vor0 = 0;
    a.div0 = n;
    a.u0 = sl.gu;
    a.v0 = sl.gv;
    a.nlev = n;
    if (launch_inverse(tb, mode, true, a, s)) return EXO_ERR_LAUNCH;
  }

  // gp = exp(gp), calcgp and the products (src/plasim.f90:2434-2446)
  exo_calcgp_fields f{};
  f.gu = slot(sl.gu);
  f.gv = slot(sl.gv);
  f.gt = slot(sl.gt);
  f.gd = slot(sl.gd);
  f.gz = slot(sl.gz);
  f.gq = slot(sl.gq);
  f.dq = slot(sl.dq);
  f.gp = slot(sl.gp);
  f.gpj = slot(sl.gpj);
  f.gpm = slot(sl.gpm);
  f.rcsq = slot(sl.rcsq);
  f.gtn = slot(sl.gtn);
  f.gqn = slot(sl.gqn);
  f.guz = slot(sl.guz);
  f.gvz = slot(sl.gvz);
  f.gphi = slot(sl.gphi);
  f.dw = slot(sl.dw);
  f.gvp = slot(sl.gvp);
  f.gut = slot(sl.gut);
  f.gvt = slot(sl.gvt);
  f.gke = slot(sl.gke);
  f.guq = nqspec ? slot(sl.guq) : nullptr;
  f.gvq = nqspec ? slot(sl.gvq) : nullptr;
  int rc = mode == MODE_F64 ? calcgp_t<double>(d, nmem, work_stride, (int)nhor, nqspec, 1, &f, s)
                            : calcgp_t<float>(d, nmem, work_stride, (int)nhor, nqspec, 1, &f, s);
  if (rc) return rc;

  // gp2fc of the nine tendency fields and gvpp, in place
  if (launch_fft(tb, mode, -1, slot(sl.a), tb.nlon, (long)(9 * n + 1) * tb.nlat, s, nmem,
                 work_stride))
    return EXO_ERR_LAUNCH;

  // scratch of the multi-block direct sums
  const size_t need = std::max(forward_scratch_bytes(tb, mode, nmem, 1),
                               dirx_scratch_bytes(tb, mode, nmem * n));
  if (need > 0 && (rc = plan_grow(&p->scratch, &p->scratch_bytes, need)) != 0) return rc;

  // fc2sp(gvpp) -> spt (field 4 nlev of the tendencies)
  {
    FwdArgs a{};
    a.in = work;
    a.sp = sp_tend;
    a.in_stride = work_stride;
    a.sp_stride = tend_stride;
    a.nmem = nmem;
    a.src0 = sl.gvp;
    a.dst0 = 4 * n;
    a.nfield = 1;
    a.table = TQC;
    if (launch_forward(tb, mode, false, a, p->scratch, p->scratch_bytes, s))
      return EXO_ERR_LAUNCH;
  }
  // mktend -> sdt, stt, szt; qtend -> sqt (src/plasim.f90:2473-2486)
  auto dirx = [&](const DirxTerm* terms, int nterm, std::initializer_list<int> ins, int out) {
    DirxArgs a{};
    int i = 0;
    for (int k : ins) a.in[i++] = slot(k);
    a.nin = i;
    a.out = (char*)sp_tend + es * 2 * tb.ncsp * out;
    a.nlev = n;
    a.nmem = nmem;
    a.in_stride = work_stride;
    a.out_stride = tend_stride;
    a.nterm = nterm;
    for (int t = 0; t < nterm; ++t) a.terms[t] = terms[t];
    return launch_dirx(tb, mode, a, p->scratch, p->scratch_bytes, s);
  };
  if (dirx(MKTEND_D, 6, {sl.gke, sl.guz, sl.gvz}, n)) return EXO_ERR_LAUNCH;
  if (dirx(MKTEND_T, 6, {sl.gvt, sl.gtn, sl.gut}, 2 * n)) return EXO_ERR_LAUNCH;
  if (dirx(MKTEND_Z, 4, {sl.gvz, sl.guz}, 0)) return EXO_ERR_LAUNCH;
  if (nqspec && dirx(QTEND_Q, 6, {sl.gqn, sl.guq, sl.gvq}, 3 * n)) return EXO_ERR_LAUNCH;
  return EXO_OK;
}
