/*
 * exo_oracle.c -- FP64 CPU restatement of PlaSim's spectral-transform core,
 * the reference that the HIP implementation in exoplasim_amd is tested against.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library; the product never does.
 *
 * Each function cites the reference routine it restates ("src/" is
 * exoplasim/plasim/src/ of the reference).  Loop order and the association
 * order of each floating-point expression follow the reference, so with
 * -ffp-contract=off the results agree with the reference compiled at
 * -fdefault-real-8 to a few ulp (checked against the compiled reference and
 * the committed golden vectors in tests/).
 *
 * Pinning (see tests/test_oracle.py):
 *   ora_inigau            numpy leggauss, golden vectors, compiled reference
 *   ora_legini            scipy lpmv closed form, quadrature orthonormality,
 *                         Q recurrence identity, compiled reference
 *   ora_sp2fc/fc2sp       round-trip identity, single-mode closed form,
 *                         compiled reference (T21 legmod, any-T pyfft copy)
 *   ora_dv2uv/uv2dv       solid-body / pure planetary vorticity closed
 *                         forms, compiled reference
 *   ora_mktend/qtend      compiled reference (T21 legmod), linearity
 *   ora_gp2fc/fc2gp       numpy rfft closed form, compiled reference fftmod
 *   ora_*_dft (fft991)    numpy rfft closed form, compiled reference fft991
 *   ora_minvers/makebm    numpy.linalg.inv
 */
#include "exo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Gaussian grid: src/gaussmod.f90                                      */
/* ------------------------------------------------------------------ */

/* normalised Legendre polynomial of degree k at p (src/gaussmod.f90:46-69) */
double ora_ql(int k, double p) {
  double z0 = acos(p);
  double z1 = 1.0, z2 = 0.0, z3 = 0.0, z4;
  for (int j = k; j >= 0; j -= 2) {
    z3 = z1 * cos(z0 * j);
    z2 = z2 + z3;
    z4 = (double)((k - j + 1) * (k + j)) * 0.5;
    z1 = z1 * z4 / (z4 + (j - 1));
  }
  if (k % 2 == 0) z2 = z2 - 0.5 * z3;
  z0 = sqrt(2.0);
  for (int j = 1; j <= k; ++j) z0 = z0 * sqrt(1.0 - 0.25 / (double)(j * j));
  return z0 * z2;
}

/* Newton correction factor (src/gaussmod.f90:75-86) */
double ora_qld(int k, double p) {
  double z = p * ora_ql(k, p) -
             sqrt((k + k + 1.0) / (k + k - 1.0)) * ora_ql(k - 1, p);
  return (p * p - 1.0) / (k * z);
}

/* Gauss-Legendre abscissas (north->south) and weights
 * (src/gaussmod.f90:5-40) */
void ora_inigau(int klat, double *pz0, double *pzw) {
  const double PI = 3.14159265358979;
  const double ZEPS = 1.0e-16;
  double z0 = PI / (2 * klat + 1);
  double z1 = 1.0 / (double)(klat * klat * 8);
  double z4 = 2.0 / (double)(klat * klat);
  for (int jlat = 1; jlat <= klat / 2; ++jlat) {
    double z2 = z0 * (2 * jlat - 0.5);
    z2 = cos(z2 + z1 / tan(z2));
    for (int jiter = 1; jiter <= 50; ++jiter) {
      double z3 = ora_ql(klat, z2) * ora_qld(klat, z2);
      z2 = z2 - z3;
      if (fabs(z3) < ZEPS) break;
    }
    double z5 = ora_ql(klat - 1, z2) / sqrt(klat - 0.5);
    pz0[jlat - 1] = z2;
    pzw[jlat - 1] = z4 * (1.0 - z2 * z2) / (z5 * z5);
    pz0[klat - jlat] = -z2;
    pzw[klat - jlat] = pzw[jlat - 1];
  }
}

/* ------------------------------------------------------------------ */
/* Legendre tables: src/legmod.f90:30-171                               */
/* ------------------------------------------------------------------ */

void ora_legini(int ntru, int nlat, const double *sid, const double *gwd,
                double *tables) {
  const int ncsp = (ntru + 1) * (ntru + 2) / 2;
  const long tsz = (long)ncsp * nlat;
  double *zpli = (double *)calloc(ncsp + 2, sizeof(double));
  double *zpld = (double *)calloc(ncsp + 2, sizeof(double));
  /* 1-based views to follow the reference index arithmetic */
  double *P = zpli - 1, *D = zpld - 1;
  for (int jlat = 0; jlat < nlat; ++jlat) {
    double zgwd = gwd[jlat];
    double zsin = sid[jlat];
    double zcsq = 1.0 - zsin * zsin;
    double zgwdcsq = zgwd / zcsq;
    double f1m = sqrt(1.5), f2m;
    P[1] = sqrt(0.5);
    P[2] = f1m * zsin;
    D[1] = 0.0;
    int lm = 2;
    for (int m = 0; m <= ntru; ++m) {
      if (m > 0) {
        lm = lm + 1;
        f2m = -f1m * sqrt(zcsq / (double)(m + m));
        f1m = f2m * sqrt(m + m + 3.0);
        P[lm] = f2m;
        if (lm < ncsp) {
          lm = lm + 1;
          P[lm] = f1m * zsin;
          D[lm - 1] = -(m * f2m * zsin);
        }
      }
      double amsq = (double)(m * m);
      for (int n = m + 2; n <= ntru; ++n) {
        lm = lm + 1;
        double z1 = sqrt(((double)((n - 1) * (n - 1)) - amsq) /
                         (double)(4 * (n - 1) * (n - 1) - 1));
        double z2 = zsin * P[lm - 1] - z1 * P[lm - 2];
        P[lm] = z2 * sqrt((double)(4 * n * n - 1) / ((double)(n * n) - amsq));
        D[lm - 1] = (1 - n) * z2 + n * z1 * P[lm - 2];
      }
      if (lm < ncsp) { /* mode (m, NTRU) */
        double z3 = sqrt(((double)(ntru * ntru) - amsq) /
                         (double)(4 * ntru * ntru - 1));
        D[lm] = -ntru * zsin * P[lm] + (ntru + ntru + 1) * P[lm - 1] * z3;
      } else { /* mode (NTRU, NTRU) */
        D[lm] = -ntru * zsin * P[lm];
      }
    }
    lm = 0;
    for (int m = 0; m <= ntru; ++m) {
      for (int n = m; n <= ntru; ++n) {
        lm = lm + 1;
        double znn1 = 0.0;
        if (n > 0) znn1 = 1.0 / (double)(n * (n + 1));
        long o = (long)jlat * ncsp + (lm - 1);
        tables[ORA_QI * tsz + o] = P[lm];
        tables[ORA_QJ * tsz + o] = D[lm];
        tables[ORA_QC * tsz + o] = P[lm] * zgwd;
        tables[ORA_QU * tsz + o] = P[lm] * znn1 * m;
        tables[ORA_QV * tsz + o] = D[lm] * znn1;
        tables[ORA_QE * tsz + o] = D[lm] * zgwdcsq;
        tables[ORA_QQ * tsz + o] = P[lm] * zgwdcsq * n * (n + 1) * 0.5;
        tables[ORA_QM * tsz + o] = P[lm] * zgwdcsq * m;
      }
    }
  }
  free(zpli);
  free(zpld);
}

/* physics filters, 1-based n as in src/legmod.f90:123-155 */
void ora_filters(int ntru, int nfilter, int ngptfilter, int nspvfilter,
                 double filterkappa, int nfilterexp, double landhoskn0,
                 double *skgpsp, double *skspgp) {
  const int ntp1 = ntru + 1;
  for (int i = 0; i < ntp1; ++i) {
    const int n = i + 1;
    double fg = 1.0, fs = 1.0, f;
    switch (nfilter) {
    case 1:
      f = 1.0 - (double)n / ntp1;
      break;
    case 2:
      f = exp(-filterkappa * pow((double)n / ntru, nfilterexp));
      break;
    case 3: {
      double r = (double)n * (double)(n + 1) / (landhoskn0 * (landhoskn0 + 1));
      f = exp(-(r * r));
      break;
    }
    case 4: {
      double r = 1.0 - (double)n / ntp1;
      f = r * r;
      break;
    }
    default:
      f = 1.0;
      ngptfilter = nspvfilter = 0;
    }
    fg = (1 - ngptfilter) * fg + ngptfilter * f;
    fs = (1 - nspvfilter) * fs + nspvfilter * f;
    skgpsp[i] = fg;
    skspgp[i] = fs;
  }
}

/* ------------------------------------------------------------------ */
/* Legendre transforms: src/legmod.f90                                  */
/* ------------------------------------------------------------------ */

#define NCSP_OF(t) (((t) + 1) * ((t) + 2) / 2)

/* inverse Legendre, unfolded over all latitudes (src/legmod.f90:230-255);
 * also sp2fcdmu (src/legmod.f90:262-287) when called with qj */
void ora_sp2fc(int ntru, int nlat, int nlon, const double *qi,
               const double *skspgp, const double *sp, double *fc) {
  const int ncsp = NCSP_OF(ntru), nh = nlon / 2;
  memset(fc, 0, sizeof(double) * (size_t)nlat * nlon);
  for (int l = 0; l < nlat; ++l) {
    const double *q = qi + (long)l * ncsp;
    double *f = fc + (long)l * nlon;
    int w = 0;
    for (int m = 0; m <= ntru; ++m) {
      double re = f[2 * m], im = f[2 * m + 1];
      for (int n = m; n <= ntru; ++n, ++w) {
        re = re + q[w] * sp[2 * w] * skspgp[n];
        im = im + q[w] * sp[2 * w + 1] * skspgp[n];
      }
      f[2 * m] = re;
      f[2 * m + 1] = im;
    }
    (void)nh;
  }
}

/* direct Legendre, single-process folded branch (src/legmod.f90:203-219) */
void ora_fc2sp(int ntru, int nlat, int nlon, const double *qc,
               const double *skgpsp, const double *fc, double *sp) {
  const int ncsp = NCSP_OF(ntru);
  memset(sp, 0, sizeof(double) * 2 * (size_t)ncsp);
  for (int l = 0; l < nlat / 2; ++l) {
    const double *q = qc + (long)l * ncsp;
    const double *fn = fc + (long)l * nlon;
    const double *fs = fc + (long)(nlat - 1 - l) * nlon;
    int w = 0;
    for (int m = 0; m <= ntru; ++m) {
      for (int n = m; n <= ntru; ++n, ++w) {
        if (((n - m) & 1) == 0) {
          sp[2 * w] = sp[2 * w] + q[w] * (fn[2 * m] + fs[2 * m]) * skgpsp[n];
          sp[2 * w + 1] =
              sp[2 * w + 1] + q[w] * (fn[2 * m + 1] + fs[2 * m + 1]) * skgpsp[n];
        } else {
          sp[2 * w] = sp[2 * w] + q[w] * (fn[2 * m] - fs[2 * m]) * skgpsp[n];
          sp[2 * w + 1] =
              sp[2 * w + 1] + q[w] * (fn[2 * m + 1] - fs[2 * m + 1]) * skgpsp[n];
        }
      }
    }
  }
}

/* direct Legendre, multi-process unfolded branch over nlpp latitudes
 * (src/legmod.f90:191-202): partial sums that mpsumsc reduces */
void ora_fc2sp_unfolded(int ntru, int nlpp, int nlon, const double *qc,
                        const double *skgpsp, const double *fc, double *sp) {
  const int ncsp = NCSP_OF(ntru);
  memset(sp, 0, sizeof(double) * 2 * (size_t)ncsp);
  for (int l = 0; l < nlpp; ++l) {
    const double *q = qc + (long)l * ncsp;
    const double *f = fc + (long)l * nlon;
    int w = 0;
    for (int m = 0; m <= ntru; ++m)
      for (int n = m; n <= ntru; ++n, ++w) {
        sp[2 * w] = sp[2 * w] + q[w] * f[2 * m] * skgpsp[n];
        sp[2 * w + 1] = sp[2 * w + 1] + q[w] * f[2 * m + 1] * skgpsp[n];
      }
  }
}

/* divergence + vorticity -> u*cos(phi), v*cos(phi) Fourier coefficients
 * (src/legmod.f90:312-349).  The reference subtracts plavor from the
 * (m=0,n=1) real coefficient of pz in place and restores it; here the
 * input is read-only and the shifted value is used directly. */
void ora_dv2uv(int ntru, int nlat, int nlon, int nlev, const double *qu,
               const double *qv, const double *skspgp, double plavor,
               const double *pd, const double *pz, double *pu, double *pv) {
  const int ncsp = NCSP_OF(ntru);
  const long nsp = 2L * ncsp, nhor = (long)nlat * nlon;
  memset(pu, 0, sizeof(double) * nhor * nlev);
  memset(pv, 0, sizeof(double) * nhor * nlev);
  double *z = (double *)malloc(sizeof(double) * nsp);
  for (int v = 0; v < nlev; ++v) {
    const double *d = pd + v * nsp;
    memcpy(z, pz + v * nsp, sizeof(double) * nsp);
    z[2] = z[2] - plavor;
    for (int l = 0; l < nlat; ++l) {
      const double *u_ = qu + (long)l * ncsp, *v_ = qv + (long)l * ncsp;
      double *fu = pu + v * nhor + (long)l * nlon;
      double *fv = pv + v * nhor + (long)l * nlon;
      int w = 0;
      for (int m = 0; m <= ntru; ++m)
        for (int n = m; n <= ntru; ++n, ++w) {
          const double s = skspgp[n];
          fu[2 * m] = fu[2 * m] + v_[w] * z[2 * w] * s + u_[w] * d[2 * w + 1] * s;
          fu[2 * m + 1] =
              fu[2 * m + 1] + v_[w] * z[2 * w + 1] * s - u_[w] * d[2 * w] * s;
          fv[2 * m] = fv[2 * m] + u_[w] * z[2 * w + 1] * s - v_[w] * d[2 * w] * s;
          fv[2 * m + 1] =
              fv[2 * m + 1] - u_[w] * z[2 * w] * s - v_[w] * d[2 * w + 1] * s;
        }
    }
  }
  free(z);
}

/* u,v Fourier coefficients -> divergence + vorticity, folded branch
 * (src/legmod.f90:391-422) */
void ora_uv2dv(int ntru, int nlat, int nlon, int nlev, const double *qe,
               const double *qm, const double *skgpsp, const double *pu,
               const double *pv, double *pd, double *pz) {
  const int ncsp = NCSP_OF(ntru);
  const long nsp = 2L * ncsp, nhor = (long)nlat * nlon;
  memset(pd, 0, sizeof(double) * nsp * nlev);
  memset(pz, 0, sizeof(double) * nsp * nlev);
  for (int v = 0; v < nlev; ++v) {
    double *d = pd + v * nsp, *z = pz + v * nsp;
    for (int l = 0; l < nlat / 2; ++l) {
      const int k = nlat - 1 - l;
      const double *e = qe + (long)l * ncsp, *mm = qm + (long)l * ncsp;
      const double *ul = pu + v * nhor + (long)l * nlon;
      const double *uk = pu + v * nhor + (long)k * nlon;
      const double *vl = pv + v * nhor + (long)l * nlon;
      const double *vk = pv + v * nhor + (long)k * nlon;
      int w = 0;
      for (int m = 0; m <= ntru; ++m)
        for (int n = m; n <= ntru; ++n, ++w) {
          const double s = skgpsp[n];
          const double ur = ul[2 * m], ui = ul[2 * m + 1];
          const double urk = uk[2 * m], uik = uk[2 * m + 1];
          const double vr = vl[2 * m], vi = vl[2 * m + 1];
          const double vrk = vk[2 * m], vik = vk[2 * m + 1];
          if (((n - m) & 1) == 0) {
            z[2 * w] = z[2 * w] + e[w] * (ur - urk) * s - mm[w] * (vi + vik) * s;
            z[2 * w + 1] =
                z[2 * w + 1] + e[w] * (ui - uik) * s + mm[w] * (vr + vrk) * s;
            d[2 * w] = d[2 * w] - e[w] * (vr - vrk) * s - mm[w] * (ui + uik) * s;
            d[2 * w + 1] =
                d[2 * w + 1] - e[w] * (vi - vik) * s + mm[w] * (ur + urk) * s;
          } else {
            z[2 * w] = z[2 * w] + e[w] * (ur + urk) * s - mm[w] * (vi - vik) * s;
            z[2 * w + 1] =
                z[2 * w + 1] + e[w] * (ui + uik) * s + mm[w] * (vr - vrk) * s;
            d[2 * w] = d[2 * w] - e[w] * (vr + vrk) * s - mm[w] * (ui - uik) * s;
            d[2 * w + 1] =
                d[2 * w + 1] - e[w] * (vi + vik) * s + mm[w] * (ur - urk) * s;
          }
        }
    }
  }
}

/* tracer tendency, folded branch (src/legmod.f90:468-496) */
void ora_qtend(int ntru, int nlat, int nlon, int nlev, const double *qc,
               const double *qe, const double *qm, const double *skgpsp,
               const double *qn, const double *uq, const double *vq,
               double *q) {
  const int ncsp = NCSP_OF(ntru);
  const long nsp = 2L * ncsp, nhor = (long)nlat * nlon;
  memset(q, 0, sizeof(double) * nsp * nlev);
  for (int v = 0; v < nlev; ++v) {
    double *o = q + v * nsp;
    for (int l = 0; l < nlat / 2; ++l) {
      const int k = nlat - 1 - l;
      const double *c = qc + (long)l * ncsp, *e = qe + (long)l * ncsp,
                   *mm = qm + (long)l * ncsp;
      const double *nl = qn + v * nhor + (long)l * nlon,
                   *nk = qn + v * nhor + (long)k * nlon;
      const double *ul = uq + v * nhor + (long)l * nlon,
                   *uk = uq + v * nhor + (long)k * nlon;
      const double *vl = vq + v * nhor + (long)l * nlon,
                   *vk = vq + v * nhor + (long)k * nlon;
      int w = 0;
      for (int m = 0; m <= ntru; ++m)
        for (int n = m; n <= ntru; ++n, ++w) {
          const double s = skgpsp[n];
          const int r = 2 * m, i = 2 * m + 1;
          if (((n - m) & 1) == 0) {
            o[2 * w] = o[2 * w] + e[w] * (vl[r] - vk[r]) * s +
                       c[w] * (nl[r] + nk[r]) * s + mm[w] * (ul[i] + uk[i]) * s;
            o[2 * w + 1] = o[2 * w + 1] + e[w] * (vl[i] - vk[i]) * s +
                           c[w] * (nl[i] + nk[i]) * s -
                           mm[w] * (ul[r] + uk[r]) * s;
          } else {
            o[2 * w] = o[2 * w] + e[w] * (vl[r] + vk[r]) * s +
                       c[w] * (nl[r] - nk[r]) * s + mm[w] * (ul[i] - uk[i]) * s;
            o[2 * w + 1] = o[2 * w + 1] + e[w] * (vl[i] + vk[i]) * s +
                           c[w] * (nl[i] - nk[i]) * s -
                           mm[w] * (ul[r] - uk[r]) * s;
          }
        }
    }
  }
}

/* fused direct transform of the adiabatic tendencies, folded branch
 * (src/legmod.f90:555-602); the caller passes fu<-gvz, fv<-guz
 * (src/plasim.f90:2478) */
void ora_mktend(int ntru, int nlat, int nlon, int nlev, const double *qc,
                const double *qe, const double *qm, const double *qq,
                const double *skgpsp, const double *tn, const double *fu,
                const double *fv, const double *ke, const double *ut,
                const double *vt, double *d, double *t, double *z) {
  const int ncsp = NCSP_OF(ntru);
  const long nsp = 2L * ncsp, nhor = (long)nlat * nlon;
  memset(d, 0, sizeof(double) * nsp * nlev);
  memset(t, 0, sizeof(double) * nsp * nlev);
  memset(z, 0, sizeof(double) * nsp * nlev);
  for (int v = 0; v < nlev; ++v) {
    double *od = d + v * nsp, *ot = t + v * nsp, *oz = z + v * nsp;
    for (int l = 0; l < nlat / 2; ++l) {
      const int k = nlat - 1 - l;
      const double *c = qc + (long)l * ncsp, *e = qe + (long)l * ncsp,
                   *mm = qm + (long)l * ncsp, *qqq = qq + (long)l * ncsp;
#define ROWS(p) const double *p##l = p + v * nhor + (long)l * nlon, \
                             *p##k = p + v * nhor + (long)k * nlon
      ROWS(tn); ROWS(fu); ROWS(fv); ROWS(ke); ROWS(ut); ROWS(vt);
#undef ROWS
      int w = 0;
      for (int m = 0; m <= ntru; ++m)
        for (int n = m; n <= ntru; ++n, ++w) {
          const double s = skgpsp[n];
          const int r = 2 * m, i = 2 * m + 1;
          /* sg = +1 for the (N+S) combination of a symmetric mode */
          const int sym = (((n - m) & 1) == 0);
#define P(a, x) (sym ? (a##l[x] + a##k[x]) : (a##l[x] - a##k[x]))
#define M(a, x) (sym ? (a##l[x] - a##k[x]) : (a##l[x] + a##k[x]))
          od[2 * w] = od[2 * w] + qqq[w] * P(ke, r) * s - e[w] * M(fv, r) * s -
                      mm[w] * P(fu, i) * s;
          od[2 * w + 1] = od[2 * w + 1] + qqq[w] * P(ke, i) * s -
                          e[w] * M(fv, i) * s + mm[w] * P(fu, r) * s;
          ot[2 * w] = ot[2 * w] + e[w] * M(vt, r) * s + c[w] * P(tn, r) * s +
                      mm[w] * P(ut, i) * s;
          ot[2 * w + 1] = ot[2 * w + 1] + e[w] * M(vt, i) * s +
                          c[w] * P(tn, i) * s - mm[w] * P(ut, r) * s;
          oz[2 * w] = oz[2 * w] + e[w] * M(fu, r) * s - mm[w] * P(fv, i) * s;
          oz[2 * w + 1] =
              oz[2 * w + 1] + e[w] * M(fu, i) * s + mm[w] * P(fv, r) * s;
#undef P
#undef M
        }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Real FFT: src/fftmod.f90 (self-sorting radix 8/4/3/2 passes)        */
/* ------------------------------------------------------------------ */

static const int kFftAllowed[] = {8,   16,  32,  48,   64,   96,  128,
                                  256, 384, 512, 1024, 2048, 4096};

int ora_fft_supported(int n) {
  for (unsigned j = 0; j < sizeof(kFftAllowed) / sizeof(int); ++j)
    if (kFftAllowed[j] == n) return 1;
  return 0;
}

/* trigs(2k+1)=cos(k*del), trigs(2k+2)=sin(k*del) (src/fftmod.f90:115-120) */
static void fft_trigs(int n, double *trigs) {
  double del = 4.0 * asin(1.0) / n;
  for (int k = 0; k <= n / 2 - 1; ++k) {
    double angle = k * del;
    trigs[2 * k] = cos(angle);
    trigs[2 * k + 1] = sin(angle);
  }
}

/* 1-based accessors over column-major (n, lot) arrays */
#define A1(i, l) a[(long)((l) - 1) * n + (i) - 1]
#define C1(i, l) c[(long)((l) - 1) * n + (i) - 1]
#define T1(i) trigs[(i) - 1]

/* first forward pass, radix 8 with 1/n scaling (src/fftmod.f90:333-373) */
static void dfft8(double *a, int n, int lot) {
  const int la = n / 8;
  const double z = 1.0 / n;
  const double zsin45 = z * sqrt(0.5);
  for (long i = 0; i < (long)la * lot; ++i) {
    long i0 = (i / la) * n + (i % la);
    long i1 = i0 + la, i2 = i1 + la, i3 = i2 + la, i4 = i3 + la,
         i5 = i4 + la, i6 = i5 + la, i7 = i6 + la;
    double a0p4 = a[i0] + a[i4];
    double a1p5 = a[i1] + a[i5];
    double a2p6 = a[i2] + a[i6];
    double a3p7 = a[i3] + a[i7];
    double a5m1 = a[i5] - a[i1];
    double a7m3 = a[i7] - a[i3];
    double a0m4 = (a[i0] - a[i4]) * z;
    double a6m2 = (a[i6] - a[i2]) * z;
    double a0p4p2p6 = a0p4 + a2p6;
    double a1p5p3p7 = a1p5 + a3p7;
    double a7m3p5m1 = (a7m3 + a5m1) * zsin45;
    double a7m3m5m1 = (a7m3 - a5m1) * zsin45;
    a[i0] = z * (a0p4p2p6 + a1p5p3p7);
    a[i7] = z * (a0p4p2p6 - a1p5p3p7);
    a[i3] = z * (a0p4 - a2p6);
    a[i4] = z * (a3p7 - a1p5);
    a[i1] = a0m4 + a7m3m5m1;
    a[i5] = a0m4 - a7m3m5m1;
    a[i2] = a7m3p5m1 + a6m2;
    a[i6] = a7m3p5m1 - a6m2;
  }
}

/* forward radix-4 pass (src/fftmod.f90:212-327); la is divided by 4 */
static void dfft4(double *a, double *c, const double *trigs, int n, int lot,
                  int *pla) {
  int la = *pla / 4;
  *pla = la;
  int i1 = la, i2 = la + i1, i3 = la + i2, i4 = la + i3, i5 = la + i4,
      i6 = la + i5, i7 = la + i6;
  int j1 = n / 2 - la, j2 = n - la, j3 = j1, j5 = j1 + la;
  for (int i = 1; i <= la; ++i)
    for (int l = 1; l <= lot; ++l) {
      double a0p2 = A1(i, l) + A1(i2 + i, l);
      double a1p3 = A1(i1 + i, l) + A1(i3 + i, l);
      C1(i, l) = a0p2 + a1p3;
      C1(j2 + i, l) = a0p2 - a1p3;
      C1(j1 + i, l) = A1(i, l) - A1(i2 + i, l);
      C1(j5 + i, l) = A1(i3 + i, l) - A1(i1 + i, l);
    }
  int jink = 2 * la;
  int j0 = la;
  j1 = j1 + jink;
  j2 = j2 - jink;
  j3 = j3 - jink;
  int j4 = j0 + la;
  j5 = j1 + la;
  int j6 = j2 + la;
  int j7 = j3 + la;
  int ibase = 4 * la;
  for (int k = la; k <= (n - 4) / 8; k += la) {
    int kb = k + k, kc = kb + kb, kd = kc + kb;
    double c1 = T1(kb + 1), s1 = T1(kb + 2);
    double c2 = T1(kc + 1), s2 = T1(kc + 2);
    double c3 = T1(kd + 1), s3 = T1(kd + 2);
    int i = ibase + 1;
    for (int j = 1; j <= la; ++j, ++i)
      for (int l = 1; l <= lot; ++l) {
        double a1p5 = c1 * A1(i1 + i, l) + s1 * A1(i5 + i, l);
        double a2p6 = c2 * A1(i2 + i, l) + s2 * A1(i6 + i, l);
        double a3p7 = c3 * A1(i3 + i, l) + s3 * A1(i7 + i, l);
        double a5m1 = c1 * A1(i5 + i, l) - s1 * A1(i1 + i, l);
        double a6m2 = c2 * A1(i6 + i, l) - s2 * A1(i2 + i, l);
        double a7m3 = c3 * A1(i7 + i, l) - s3 * A1(i3 + i, l);
        double a0 = A1(i, l) + a2p6;
        double a2 = A1(i, l) - a2p6;
        double a1 = a1p5 + a3p7;
        double a3 = a3p7 - a1p5;
        double b0 = A1(i4 + i, l) + a6m2;
        double b2 = A1(i4 + i, l) - a6m2;
        double b1 = a5m1 + a7m3;
        double b3 = a5m1 - a7m3;
        C1(j0 + j, l) = a0 + a1;
        C1(j2 + j, l) = a0 - a1;
        C1(j4 + j, l) = b0 + b1;
        C1(j6 + j, l) = b1 - b0;
        C1(j1 + j, l) = a2 + b3;
        C1(j3 + j, l) = a2 - b3;
        C1(j5 + j, l) = a3 + b2;
        C1(j7 + j, l) = a3 - b2;
      }
    ibase = ibase + 8 * la;
    j0 = j0 + jink;
    j1 = j1 + jink;
    j2 = j2 - jink;
    j3 = j3 - jink;
    j4 = j0 + la;
    j5 = j1 + la;
    j6 = j2 + la;
    j7 = j3 + la;
  }
  if (j1 <= j2) {
    const double sin45 = sqrt(0.5);
    int i = ibase + 1;
    for (int j = 1; j <= la; ++j, ++i)
      for (int l = 1; l <= lot; ++l) {
        double a1p3 = sin45 * (A1(i1 + i, l) + A1(i3 + i, l));
        double a1m3 = sin45 * (A1(i1 + i, l) - A1(i3 + i, l));
        C1(j0 + j, l) = A1(i, l) + a1m3;
        C1(j1 + j, l) = A1(i, l) - a1m3;
        C1(j4 + j, l) = -A1(i2 + i, l) - a1p3;
        C1(j5 + j, l) = A1(i2 + i, l) - a1p3;
      }
  }
  if (la == 1) {
    for (int l = 1; l <= lot; ++l) {
      A1(1, l) = C1(1, l);
      A1(2, l) = 0.0;
      for (int i = 3; i <= n; ++i) A1(i, l) = C1(i - 1, l);
    }
  } else {
    memcpy(a, c, sizeof(double) * (size_t)n * lot);
  }
}

/* last forward pass, radix 3, one row (src/fftmod.f90:161-206) */
static void dfft3(double *a, double *c, const double *trigs, int n) {
  const double SIN60 = 0.866025403784438;
  const int l = 1;
  int ja = 1, jb = 2 * (n / 3) + 1, jc = jb;
  C1(ja, l) = A1(1, l) + A1(2, l) + A1(3, l);
  C1(ja + 1, l) = 0.0;
  C1(jb, l) = A1(1, l) - 0.5 * (A1(2, l) + A1(3, l));
  C1(jb + 1, l) = SIN60 * (A1(3, l) - A1(2, l));
  ja = 3;
  jb = jb + 2;
  jc = jc - 2;
  for (int i = 4; i <= n - 8; i += 6) {
    double c1 = T1(ja), s1 = T1(ja + 1), c2 = T1(ja + ja - 1), s2 = T1(ja + ja);
    double a1 = (c1 * A1(i + 1, l) + s1 * A1(i + 4, l)) +
                (c2 * A1(i + 2, l) + s2 * A1(i + 5, l));
    double b1 = (c1 * A1(i + 4, l) - s1 * A1(i + 1, l)) +
                (c2 * A1(i + 5, l) - s2 * A1(i + 2, l));
    double a2 = A1(i, l) - 0.5 * a1;
    double b2 = A1(i + 3, l) - 0.5 * b1;
    double a3 = SIN60 * ((c1 * A1(i + 1, l) + s1 * A1(i + 4, l)) -
                         (c2 * A1(i + 2, l) + s2 * A1(i + 5, l)));
    double b3 = SIN60 * ((c1 * A1(i + 4, l) - s1 * A1(i + 1, l)) -
                         (c2 * A1(i + 5, l) - s2 * A1(i + 2, l)));
    C1(ja, l) = A1(i, l) + a1;
    C1(ja + 1, l) = A1(i + 3, l) + b1;
    C1(jb, l) = a2 + b3;
    C1(jb + 1, l) = b2 - a3;
    C1(jc, l) = a2 - b3;
    C1(jc + 1, l) = -b2 - a3;
    ja += 2;
    jb += 2;
    jc -= 2;
  }
  if (ja <= jc) {
    C1(ja, l) = A1(n - 2, l) + 0.5 * (A1(n - 1, l) - A1(n, l));
    C1(ja + 1, l) = -SIN60 * (A1(n - 1, l) + A1(n, l));
  }
  memcpy(a, c, sizeof(double) * n);
}

/* last forward pass, radix 2, one row (src/fftmod.f90:128-155) */
static void dfft2(double *a, double *c, const double *trigs, int n) {
  const int l = 1;
  C1(1, l) = A1(1, l) + A1(2, l);
  C1(2, l) = 0.0;
  int ja = 3, jb = n - 1;
  for (int i = 3; i <= n - 5; i += 4) {
    double c1 = T1(ja), s1 = T1(ja + 1);
    double a1p3 = c1 * A1(i + 1, l) + s1 * A1(i + 3, l);
    double a3m1 = c1 * A1(i + 3, l) - s1 * A1(i + 1, l);
    C1(ja, l) = A1(i, l) + a1p3;
    C1(jb, l) = A1(i, l) - a1p3;
    C1(ja + 1, l) = a3m1 + A1(i + 2, l);
    C1(jb + 1, l) = a3m1 - A1(i + 2, l);
    ja += 2;
    jb -= 2;
  }
  C1(ja, l) = A1(n - 1, l);
  C1(ja + 1, l) = -A1(n, l);
  memcpy(a, c, sizeof(double) * n);
}

/* inverse radix-4 pass (src/fftmod.f90:379-486); the data argument is
 * `d` and `a` is the scratch copy, as in the reference where the dummy
 * argument is named c; la is multiplied by 4 */
static void ifft4(double *d, double *a, const double *trigs, int n, int lot,
                  int *pla) {
  const int la = *pla;
  double *c = d; /* so that C1() addresses the data argument */
  if (la == 1) {
    for (int l = 1; l <= lot; ++l) {
      A1(1, l) = 0.5 * C1(1, l);
      A1(n, l) = 0.0;
      for (int i = 2; i <= n - 1; ++i) A1(i, l) = C1(i + 1, l);
    }
  } else {
    memcpy(a, c, sizeof(double) * (size_t)n * lot);
  }
  const int kstop = (n - 4) / 8;
  int i1 = n / 2 - la, i2 = n - la, i5 = i1 + la;
  int j1 = la, j2 = la + j1, j3 = la + j2, j4 = la + j3, j5 = la + j4,
      j6 = la + j5, j7 = la + j6;
  for (int i = 1; i <= la; ++i)
    for (int l = 1; l <= lot; ++l) {
      C1(i, l) = A1(i, l) + A1(i2 + i, l) + A1(i1 + i, l);
      C1(j1 + i, l) = A1(i, l) - A1(i2 + i, l) - A1(i5 + i, l);
      C1(j2 + i, l) = A1(i, l) + A1(i2 + i, l) - A1(i1 + i, l);
      C1(j3 + i, l) = A1(i, l) - A1(i2 + i, l) + A1(i5 + i, l);
    }
  const int iink = 2 * la;
  int jbase = 4 * la + 1;
  int i0 = la;
  i1 = i0 + n / 2;
  i2 = n - 3 * la;
  int i3 = i2 - n / 2;
  int i4 = i0 + la;
  i5 = i1 + la;
  int i6 = i2 + la;
  int i7 = i3 + la;
  for (int k = la; k <= kstop; k += la) {
    int kb = k + k, kc = kb + kb, kd = kc + kb;
    double c1 = T1(kb + 1), s1 = T1(kb + 2);
    double c2 = T1(kc + 1), s2 = T1(kc + 2);
    double c3 = T1(kd + 1), s3 = T1(kd + 2);
    for (int i = 1; i <= la; ++i) {
      const int j = jbase;
      for (int l = 1; l <= lot; ++l) {
        double a0p2 = A1(i0 + i, l) + A1(i2 + i, l);
        double a0m2 = A1(i0 + i, l) - A1(i2 + i, l);
        double a1p3 = A1(i1 + i, l) + A1(i3 + i, l);
        double a1m3 = A1(i1 + i, l) - A1(i3 + i, l);
        double a4p6 = A1(i4 + i, l) + A1(i6 + i, l);
        double a4m6 = A1(i4 + i, l) - A1(i6 + i, l);
        double a5p7 = A1(i5 + i, l) + A1(i7 + i, l);
        double a5m7 = A1(i5 + i, l) - A1(i7 + i, l);
        double a0p2m1p3 = a0p2 - a1p3;
        double a4m6m5m7 = a4m6 - a5m7;
        C1(j, l) = a0p2 + a1p3;
        C1(j4 + j, l) = a4m6 + a5m7;
        C1(j2 + j, l) = c2 * a0p2m1p3 - s2 * a4m6m5m7;
        C1(j6 + j, l) = s2 * a0p2m1p3 + c2 * a4m6m5m7;
        C1(j1 + j, l) = c1 * (a0m2 - a5p7) - s1 * (a4p6 + a1m3);
        C1(j5 + j, l) = s1 * (a0m2 - a5p7) + c1 * (a4p6 + a1m3);
        C1(j3 + j, l) = c3 * (a0m2 + a5p7) - s3 * (a4p6 - a1m3);
        C1(j7 + j, l) = s3 * (a0m2 + a5p7) + c3 * (a4p6 - a1m3);
      }
      jbase = jbase + 1;
    }
    i0 += iink;
    i1 += iink;
    i2 -= iink;
    i3 -= iink;
    i4 += iink;
    i5 += iink;
    i6 -= iink;
    i7 -= iink;
    jbase = jbase + 7 * la;
  }
  if (i1 <= i2) {
    const double sin45 = sqrt(0.5);
    for (int i = 1; i <= la; ++i) {
      const int j = jbase;
      for (int l = 1; l <= lot; ++l) {
        C1(j, l) = A1(i0 + i, l) + A1(i1 + i, l);
        C1(j1 + j, l) = sin45 * ((A1(i0 + i, l) - A1(i1 + i, l)) -
                                 (A1(la + i0 + i, l) + A1(la + i1 + i, l)));
        C1(j2 + j, l) = A1(la + i1 + i, l) - A1(la + i0 + i, l);
        C1(j3 + j, l) = -(sin45 * ((A1(i0 + i, l) - A1(i1 + i, l)) +
                                   (A1(la + i0 + i, l) + A1(la + i1 + i, l))));
      }
      jbase = jbase + 1;
    }
  }
  *pla = la * 4;
}

/* first inverse pass, radix 2 (src/fftmod.f90:492-521) */
static void ifft2(double *a, double *c, const double *trigs, int n, int lot,
                  int *pla) {
  for (int l = 1; l <= lot; ++l) {
    C1(1, l) = 0.5 * A1(1, l);
    C1(2, l) = C1(1, l);
  }
  int ia = 3, ib = n - 1;
  for (int j = 3; j <= n - 5; j += 4) {
    double c1 = T1(ia), s1 = T1(ia + 1);
    for (int l = 1; l <= lot; ++l) {
      double amb = A1(ia, l) - A1(ib, l);
      double apb = A1(ia + 1, l) + A1(ib + 1, l);
      C1(j, l) = A1(ia, l) + A1(ib, l);
      C1(j + 2, l) = A1(ia + 1, l) - A1(ib + 1, l);
      C1(j + 1, l) = c1 * amb - s1 * apb;
      C1(j + 3, l) = s1 * amb + c1 * apb;
    }
    ia += 2;
    ib -= 2;
  }
  for (int l = 1; l <= lot; ++l) {
    C1(n - 1, l) = A1(ia, l);
    C1(n, l) = -A1(ia + 1, l);
  }
  memcpy(a, c, sizeof(double) * (size_t)n * lot);
  *pla = 2;
}

/* first inverse pass, radix 3 (src/fftmod.f90:527-572) */
static void ifft3(double *a, double *c, const double *trigs, int n, int lot,
                  int *pla) {
  const double SIN60 = 0.866025403784438;
  int ib = 2 * (n / 3) + 1;
  for (int l = 1; l <= lot; ++l) {
    C1(1, l) = 0.5 * A1(1, l) + A1(ib, l);
    C1(2, l) = 0.5 * A1(1, l) - 0.5 * A1(ib, l) - SIN60 * A1(ib + 1, l);
    C1(3, l) = 0.5 * A1(1, l) - 0.5 * A1(ib, l) + SIN60 * A1(ib + 1, l);
  }
  int ia = 3, ic = ib - 2;
  ib = ib + 2;
  for (int j = 4; j <= n - 8; j += 6) {
    double c1 = T1(ia), s1 = T1(ia + 1), c2 = T1(ia + ia - 1), s2 = T1(ia + ia);
    for (int l = 1; l <= lot; ++l) {
      double hbpc = A1(ia, l) - 0.5 * (A1(ib, l) + A1(ic, l));
      double hbmc = A1(ia + 1, l) - 0.5 * (A1(ib + 1, l) - A1(ic + 1, l));
      double sbmc = SIN60 * (A1(ib, l) - A1(ic, l));
      double sbpc = SIN60 * (A1(ib + 1, l) + A1(ic + 1, l));
      C1(j, l) = A1(ia, l) + A1(ib, l) + A1(ic, l);
      C1(j + 3, l) = A1(ia + 1, l) + A1(ib + 1, l) - A1(ic + 1, l);
      C1(j + 1, l) = c1 * (hbpc - sbpc) - s1 * (hbmc + sbmc);
      C1(j + 4, l) = s1 * (hbpc - sbpc) + c1 * (hbmc + sbmc);
      C1(j + 2, l) = c2 * (hbpc + sbpc) - s2 * (hbmc - sbmc);
      C1(j + 5, l) = s2 * (hbpc + sbpc) + c2 * (hbmc - sbmc);
    }
    ia += 2;
    ib += 2;
    ic -= 2;
  }
  for (int l = 1; l <= lot; ++l) {
    C1(n - 2, l) = A1(ia, l);
    C1(n - 1, l) = 0.5 * A1(ia, l) - SIN60 * A1(ia + 1, l);
    C1(n, l) = -(0.5 * A1(ia, l)) - SIN60 * A1(ia + 1, l);
  }
  memcpy(a, c, sizeof(double) * (size_t)n * lot);
  *pla = 3;
}

/* last inverse pass, radix 8 (src/fftmod.f90:578-618) */
static void ifft8(double *a, int n, int lot) {
  const double SQRT2 = 1.414213562373095;
  const int la = n / 8;
  for (long i = 0; i < (long)la * lot; ++i) {
    long i0 = (i / la) * n + (i % la);
    long i1 = i0 + la, i2 = i1 + la, i3 = i2 + la, i4 = i3 + la,
         i5 = i4 + la, i6 = i5 + la, i7 = i6 + la;
    double a0p7 = a[i0] + a[i7];
    double a0m7 = a[i0] - a[i7];
    double a1p5 = a[i1] + a[i5];
    double a1m5 = a[i1] - a[i5];
    double a2p6 = a[i2] + a[i6];
    double a2m6 = a[i2] - a[i6];
    double a0p7p3 = a0p7 + a[i3];
    double a0p7m3 = a0p7 - a[i3];
    double a0m7p4 = 2.0 * (a0m7 + a[i4]);
    double a0m7m4 = 2.0 * (a0m7 - a[i4]);
    double a1m5p2p6 = SQRT2 * (a1m5 + a2p6);
    double a1m5m2p6 = SQRT2 * (a1m5 - a2p6);
    a[i0] = 2.0 * (a0p7p3 + a1p5);
    a[i2] = 2.0 * (a0p7m3 - a2m6);
    a[i4] = 2.0 * (a0p7p3 - a1p5);
    a[i6] = 2.0 * (a0p7m3 + a2m6);
    a[i1] = a0m7m4 + a1m5m2p6;
    a[i3] = a0m7p4 - a1m5p2p6;
    a[i5] = a0m7m4 - a1m5m2p6;
    a[i7] = a0m7p4 + a1m5p2p6;
  }
}

/* gridpoint -> Fourier, in place (src/fftmod.f90:29-58) */
int ora_gp2fc(double *a, int n, int lot) {
  if (!ora_fft_supported(n)) return -1;
  double *trigs = (double *)malloc(sizeof(double) * n);
  double *c = (double *)malloc(sizeof(double) * (size_t)n * lot);
  fft_trigs(n, trigs);
  dfft8(a, n, lot);
  int la = n / 8;
  while (la >= 4) dfft4(a, c, trigs, n, lot, &la);
  if (la == 3)
    for (int l = 0; l < lot; ++l) dfft3(a + (long)l * n, c, trigs, n);
  if (la == 2)
    for (int l = 0; l < lot; ++l) dfft2(a + (long)l * n, c, trigs, n);
  free(c);
  free(trigs);
  return 0;
}

/* Fourier -> gridpoint, in place (src/fftmod.f90:64-87) */
int ora_fc2gp(double *a, int n, int lot) {
  if (!ora_fft_supported(n)) return -1;
  double *trigs = (double *)malloc(sizeof(double) * n);
  double *c = (double *)malloc(sizeof(double) * (size_t)n * lot);
  fft_trigs(n, trigs);
  int nf = n / 8;
  while (nf >= 4) nf = nf / 4;
  int la = 1;
  if (nf == 2) ifft2(a, c, trigs, n, lot, &la);
  if (nf == 3) ifft3(a, c, trigs, n, lot, &la);
  while (la < n / 8) ifft4(a, c, trigs, n, lot, &la);
  ifft8(a, n, lot);
  free(c);
  free(trigs);
  return 0;
}

#undef A1
#undef C1
#undef T1

/* ------------------------------------------------------------------ */
/* fft991mod by its documented definition (src/fft991mod.f90:148-155)   */
/* ------------------------------------------------------------------ */

/* A(k) = (1/N) sum x(j) cos(2 pi jk/N), B(k) = -(1/N) sum x(j) sin(...),
 * k = 0..N/2-1, B(0) = 0; the Nyquist coefficient is not stored
 * (gp2fc keeps a(1:n) of the n+2 work row, src/fft991mod.f90:46-50). */
void ora_gp2fc_dft(double *a, int n, int lot) {
  double *cs = (double *)malloc(sizeof(double) * 2 * n);
  double *x = (double *)malloc(sizeof(double) * n);
  for (int r = 0; r < n; ++r) {
    cs[2 * r] = cos(2.0 * M_PI * r / n);
    cs[2 * r + 1] = sin(2.0 * M_PI * r / n);
  }
  for (int l = 0; l < lot; ++l) {
    double *row = a + (long)l * n;
    memcpy(x, row, sizeof(double) * n);
    for (int k = 0; k < n / 2; ++k) {
      long double sa = 0.0L, sb = 0.0L;
      for (int j = 0; j < n; ++j) {
        int r = (int)(((long)j * k) % n);
        sa += (long double)x[j] * cs[2 * r];
        sb += (long double)x[j] * cs[2 * r + 1];
      }
      row[2 * k] = (double)(sa / n);
      row[2 * k + 1] = (k == 0) ? 0.0 : (double)(-sb / n);
    }
  }
  free(x);
  free(cs);
}

/* x(j) = A(0) + 2 sum_{k=1}^{N/2-1} (A(k) cos - B(k) sin)(2 pi jk/N);
 * B(0) and the Nyquist coefficient are not used
 * (src/fft991mod.f90:64-80, 148-152). */
void ora_fc2gp_dft(double *a, int n, int lot) {
  double *cs = (double *)malloc(sizeof(double) * 2 * n);
  double *f = (double *)malloc(sizeof(double) * n);
  for (int r = 0; r < n; ++r) {
    cs[2 * r] = cos(2.0 * M_PI * r / n);
    cs[2 * r + 1] = sin(2.0 * M_PI * r / n);
  }
  for (int l = 0; l < lot; ++l) {
    double *row = a + (long)l * n;
    memcpy(f, row, sizeof(double) * n);
    for (int j = 0; j < n; ++j) {
      long double s = 0.0L;
      for (int k = 1; k < n / 2; ++k) {
        int r = (int)(((long)j * k) % n);
        s += (long double)f[2 * k] * cs[2 * r] -
             (long double)f[2 * k + 1] * cs[2 * r + 1];
      }
      row[j] = (double)((long double)f[0] + 2.0L * s);
    }
  }
  free(f);
  free(cs);
}

/* ------------------------------------------------------------------ */
/* One round trip (sp2fc, sp2fcdmu, dv2uv -> fc2gp -> gp2fc -> fc2sp)  */
/* ------------------------------------------------------------------ */

long ora_roundtrip_worksize(int ntru, int nlat, int nlon, int nlev) {
  (void)ntru;
  (void)nlev;
  return (long)nlat * nlon * 2;
}

int ora_roundtrip(int ntru, int nlat, int nlon, int nlev, const double *tables,
                  const double *skspgp, const double *skgpsp, double plavor,
                  int use_dft, const double *sp_in, double *gp,
                  double *sp_out, double *work) {
  const int ncsp = NCSP_OF(ntru);
  const long nsp = 2L * ncsp, nhor = (long)nlat * nlon, tsz = (long)ncsp * nlat;
  const int nfl = 3 * nlev + 1;
  const double *qi = tables + ORA_QI * tsz, *qj = tables + ORA_QJ * tsz,
               *qc = tables + ORA_QC * tsz, *qu = tables + ORA_QU * tsz,
               *qv = tables + ORA_QV * tsz;
  if (!use_dft && !ora_fft_supported(nlon)) return -1;
  /* inverse: each field-level, then d/dmu of ln(ps), then u, v */
  for (int f = 0; f < nfl; ++f)
    ora_sp2fc(ntru, nlat, nlon, qi, skspgp, sp_in + f * nsp, gp + f * nhor);
  ora_sp2fc(ntru, nlat, nlon, qj, skspgp, sp_in + (nfl - 1) * nsp,
            gp + nfl * nhor);
  /* vorticity = fields [0,nlev), divergence = [nlev, 2 nlev) */
  ora_dv2uv(ntru, nlat, nlon, nlev, qu, qv, skspgp, plavor,
            sp_in + (long)nlev * nsp, sp_in, gp + (nfl + 1) * nhor,
            gp + (nfl + 1 + nlev) * nhor);
  const int ngp = nfl + 1 + 2 * nlev;
  if (use_dft)
    ora_fc2gp_dft(gp, nlon, nlat * ngp);
  else
    ora_fc2gp(gp, nlon, nlat * ngp);
  /* forward: the nfl prognostic fields */
  for (int f = 0; f < nfl; ++f) {
    memcpy(work, gp + f * nhor, sizeof(double) * nhor);
    if (use_dft)
      ora_gp2fc_dft(work, nlon, nlat);
    else
      ora_gp2fc(work, nlon, nlat);
    ora_fc2sp(ntru, nlat, nlon, qc, skgpsp, work, sp_out + f * nsp);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Semi-implicit matrices: src/plasim.f90:1683-1888                      */
/* ------------------------------------------------------------------ */

/* LU decomposition with implicit pivoting (src/plasim.f90:1849-1888);
 * a is column-major n x n */
static void ludcmp(double *a, int n, int *indx) {
#define AA(i, j) a[(long)((j) - 1) * n + (i) - 1]
  double *vv = (double *)malloc(sizeof(double) * n);
  for (int i = 1; i <= n; ++i) {
    double mx = 0.0;
    for (int j = 1; j <= n; ++j) mx = fmax(mx, fabs(AA(i, j)));
    vv[i - 1] = 1.0 / mx;
  }
  int imax = 1;
  for (int j = 1; j <= n; ++j) {
    for (int i = 2; i <= j - 1; ++i) {
      double s = 0.0;
      for (int k = 1; k <= i - 1; ++k) s += AA(i, k) * AA(k, j);
      AA(i, j) = AA(i, j) - s;
    }
    double aamax = 0.0;
    for (int i = j; i <= n; ++i) {
      if (j > 1) {
        double s = 0.0;
        for (int k = 1; k <= j - 1; ++k) s += AA(i, k) * AA(k, j);
        AA(i, j) = AA(i, j) - s;
      }
      double dum = vv[i - 1] * fabs(AA(i, j));
      if (dum >= aamax) {
        imax = i;
        aamax = dum;
      }
    }
    if (j != imax) {
      for (int k = 1; k <= n; ++k) {
        double dum = AA(imax, k);
        AA(imax, k) = AA(j, k);
        AA(j, k) = dum;
      }
      vv[imax - 1] = vv[j - 1];
    }
    indx[j - 1] = imax;
    if (AA(j, j) == 0.0) AA(j, j) = 2.2250738585072014e-308;
    if (j < n)
      for (int i = j + 1; i <= n; ++i) AA(i, j) = AA(i, j) / AA(j, j);
  }
  free(vv);
}

/* forward/back substitution (src/plasim.f90:1822-1847) */
static void lubksb(const double *a, int n, const int *indx, double *b) {
  int k = 0;
  for (int i = 1; i <= n; ++i) {
    int l = indx[i - 1];
    double sum = b[l - 1];
    b[l - 1] = b[i - 1];
    if (k > 0) {
      for (int j = k; j <= i - 1; ++j) sum = sum - AA(i, j) * b[j - 1];
    } else if (sum != 0.0) {
      k = i;
    }
    b[i - 1] = sum;
  }
  for (int i = n; i >= 1; --i) {
    double sum = b[i - 1];
    for (int j = i + 1; j <= n; ++j) sum = sum - AA(i, j) * b[j - 1];
    b[i - 1] = sum / AA(i, i);
  }
#undef AA
}

/* matrix inverse in place (src/plasim.f90:1804-1820) */
int ora_minvers(double *a, int n) {
  double *b = (double *)calloc((size_t)n * n, sizeof(double));
  int *indx = (int *)malloc(sizeof(int) * n);
  for (int j = 0; j < n; ++j) b[(long)j * n + j] = 1.0;
  ludcmp(a, n, indx);
  for (int j = 0; j < n; ++j) lubksb(a, n, indx, b + (long)j * n);
  memcpy(a, b, sizeof(double) * (size_t)n * n);
  free(b);
  free(indx);
  return 0;
}

/* bm1(:,:,jn) = inverse(delt^2 (t0 dsigma + g.tau) + I/(jn(jn+1)))
 * (src/plasim.f90:1683-1703); g, tau column-major nlev x nlev, bm1 is
 * nlev x nlev x ntru column-major */
void ora_makebm(int nlev, int ntru, double delt, const double *t0,
                const double *dsigma, const double *g, const double *tau,
                double *bm1) {
  const double zdeltsq = delt * delt;
  const long nsq = (long)nlev * nlev;
  for (int j1 = 0; j1 < nlev; ++j1)
    for (int j2 = 0; j2 < nlev; ++j2) {
      double dp = 0.0;
      for (int k = 0; k < nlev; ++k)
        dp += g[(long)j1 * nlev + k] * tau[(long)k * nlev + j2];
      const double zaq = zdeltsq * (t0[j1] * dsigma[j2] + dp);
      for (int jn = 0; jn < ntru; ++jn) bm1[jn * nsq + (long)j1 * nlev + j2] = zaq;
    }
  for (int jn = 1; jn <= ntru; ++jn) {
    double *b = bm1 + (jn - 1) * nsq;
    for (int j = 0; j < nlev; ++j)
      b[(long)j * nlev + j] = b[(long)j * nlev + j] + 1.0 / (double)(jn * (jn + 1));
    ora_minvers(b, nlev);
  }
}

/* semi-implicit step of spectrala, sections 1-3 (src/plasim.f90:2817-2876),
 * one process (NSPP = nsp): overwrites sdt with the divergence at time t,
 * updates spt and stt.  g, tau, bm1 column-major; spectral arrays
 * (nsp, nlev) column-major; nindex(jsp) = total wavenumber. */
void ora_si_apply(int nlev, int nsp, const int *nindex, double delt, const double *t0,
                  const double *dsigma, const double *g, const double *tau,
                  const double *bm1, const double *sop, const double *spm,
                  const double *atm, const double *adm, double *stt, double *sdt,
                  double *spt) {
  double *zgt = (double *)malloc(sizeof(double) * nsp * nlev);
  double *zgm = (double *)malloc(sizeof(double) * nsp * nlev);
  double *adt = (double *)malloc(sizeof(double) * nsp * nlev);
  memcpy(adt, sdt, sizeof(double) * nsp * nlev);
  /* 1.a */
  for (int jlev = 0; jlev < nlev; ++jlev)
    for (int jsp = 0; jsp < nsp; ++jsp) {
      double a = 0.0, b = 0.0;
      for (int k = 0; k < nlev; ++k) {
        a += g[jlev * nlev + k] * stt[k * nsp + jsp];
        b += g[jlev * nlev + k] * atm[k * nsp + jsp];
      }
      zgt[jlev * nsp + jsp] = a;
      zgm[jlev * nsp + jsp] = b;
    }
  /* 1.b */
  for (int jlev = 0; jlev < nlev; ++jlev)
    for (int jsp = 0; jsp < nsp; ++jsp) {
      const int jn = nindex[jsp];
      double zsum = 0.0;
      if (jn > 0) {
        const double zq = 1.0 / (double)(jn * (jn + 1));
        for (int j2 = 0; j2 < nlev; ++j2) {
          const double z0 = t0[j2];
          const double zt = zgt[j2 * nsp + jsp] - z0 * spt[jsp];
          const double zm = zgm[j2 * nsp + jsp] + z0 * spm[jsp];
          const double za = adt[j2 * nsp + jsp] * zq + sop[jsp];
          const double zb = adm[j2 * nsp + jsp] * zq;
          const double zz = zb + delt * (zm + za + delt * zt);
          zsum = zsum + zz * bm1[(long)(jn - 1) * nlev * nlev + jlev * nlev + j2];
        }
      }
      sdt[jlev * nsp + jsp] = zsum;
    }
  for (int jlev = 0; jlev < nlev; ++jlev) sdt[jlev * nsp] = sdt[jlev * nsp + 1] = 0.0;
  /* 2 */
  for (int jlev = 0; jlev < nlev; ++jlev)
    for (int jsp = 0; jsp < nsp; ++jsp) spt[jsp] = spt[jsp] + dsigma[jlev] * sdt[jlev * nsp + jsp];
  /* 3 */
  for (int jlev = 0; jlev < nlev; ++jlev)
    for (int jsp = 0; jsp < nsp; ++jsp) {
      double s = 0.0;
      for (int k = 0; k < nlev; ++k) s += tau[jlev * nlev + k] * sdt[k * nsp + jsp];
      stt[jlev * nsp + jsp] = stt[jlev * nsp + jsp] - s;
    }
  free(zgt);
  free(zgm);
  free(adt);
}
