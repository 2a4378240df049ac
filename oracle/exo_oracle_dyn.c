/*
 * exo_oracle_dyn.c -- FP64 CPU restatement of the grid-point dynamics that
 * sit between PlaSim's inverse and direct transforms: the vertical
 * matrices of the semi-implicit scheme (initsi) and the nonlinear
 * grid-point tendencies (calcgp).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ and the checks of bench.py load
 * this library; the product never does.
 *
 * Loop order and association of every floating-point expression follow the
 * reference ("src/" = exoplasim/plasim/src/); with -ffp-contract=off the
 * results agree with the reference compiled at -fdefault-real-8 to rounding
 * (tests/test_oracle.py checks both against the compiled reference and the
 * golden vectors written from it).
 *
 * Layouts: vertical matrices (nlev, nlev) column-major as the Fortran
 * arrays, m(i, j) = m[j * nlev + i]; grid fields [lev][hor].
 */
#include <math.h>
#include <stdlib.h>

#include "exo_oracle.h"

/* initsi (src/plasim.f90:1723-1770): tkp, t01s2, g, c, tau from the
 * half-level sigmas, layer thicknesses, rdsig, t0 and kappa */
void ora_initsi(int nlev, const double *sigmah, const double *dsigma, const double *rdsig,
                const double *t0, double akap, double *g, double *tau, double *c,
                double *t01s2, double *tkp) {
  const int n = nlev;
#define M(a, i, j) a[(long)(j) * n + (i)]
  double *zalp = (double *)calloc((size_t)n, sizeof(double));
  for (int j = 0; j < n; ++j) tkp[j] = akap * t0[j];
  for (int j = 0; j < n - 1; ++j) t01s2[j] = t0[j + 1] - t0[j];
  t01s2[n - 1] = 0.0;
  for (int j = 1; j < n; ++j) zalp[j] = log(sigmah[j]) - log(sigmah[j - 1]);
  for (long k = 0; k < (long)n * n; ++k) g[k] = 0.0;
  M(g, 0, 0) = 1.0;
  for (int jl = 1; jl < n; ++jl) {
    M(g, jl, jl) = 1.0 - zalp[jl] * sigmah[jl - 1] / dsigma[jl];
    for (int k = 0; k < jl; ++k) M(g, jl, k) = zalp[jl];
  }
  for (int jl = 0; jl < n; ++jl)
    for (int k = 0; k < n; ++k) M(c, jl, k) = M(g, k, jl) * (dsigma[jl] / dsigma[k]);
  double zt01s2 = t01s2[0];
  double zsig = sigmah[0];
  M(tau, 0, 0) = 0.5 * zt01s2 * (zsig - 1.0) + tkp[0];
  for (int k = 1; k < n; ++k) M(tau, k, 0) = 0.5 * zt01s2 * dsigma[k];
  for (int jl = 1; jl < n; ++jl) {
    const double zttm = zt01s2, zsigm = zsig;
    zt01s2 = t01s2[jl];
    zsig = sigmah[jl];
    for (int j2 = 0; j2 < n; ++j2) {
      const double ztm = (j2 <= jl) ? 1.0 : 0.0;
      const double ztmm = (j2 < jl) ? 1.0 : 0.0;
      double ztau = zttm * (zsigm - ztmm);
      if (jl < n - 1) ztau = ztau + zt01s2 * (zsig - ztm);
      ztau = ztau * rdsig[jl] * dsigma[j2];
      if (j2 <= jl) ztau = ztau + tkp[jl] * M(c, j2, jl);
      M(tau, j2, jl) = ztau;
    }
  }
#undef M
  free(zalp);
}

/* calcgp (src/plasim.f90:2587-2750), one grid point h after another.
 * In: gu gv gt gd gz gq [lev][hor], gp (= ps / psurf), gpj, gpm (the
 * Fourier-space d/dlambda of ln ps after fc2gp), rcsq [hor], dq [lev][hor]
 * (used when nqspec == 0), the vertical set-up (sigmah, dsigma, rdsig, t0,
 * t01s2, tkp, c, g) and scal = (akap, rdbrv, adv, psurf, ww).
 * Out: gtn gqn guz gvz gphi dw [lev][hor] and gvp [hor]. */
void ora_calcgp(int nlev, long nhor, const double *gu, const double *gv, const double *gt,
                const double *gd, const double *gz, const double *gq, const double *gp,
                const double *gpj, const double *gpm, const double *dq, const double *rcsq,
                const double *sigmah, const double *dsigma, const double *rdsig,
                const double *t0, const double *t01s2, const double *tkp, const double *c,
                const double *g, const double *scal, int nqspec, double *gtn, double *gqn,
                double *guz, double *gvz, double *gvp, double *gphi, double *dw) {
  const int n = nlev;
  const double akap = scal[0], rdbrv = scal[1], adv = scal[2], psurf = scal[3], ww = scal[4];
  double *zv = (double *)malloc(sizeof(double) * 10 * n);
  double *zvgpg = zv, *ztv1 = zv + n, *ztv2 = zv + 2 * n, *zsdotp = zv + 3 * n;
  double *gtd = zv + 4 * n, *gqm = zv + 5 * n, *gud = zv + 6 * n, *gvd = zv + 7 * n;
#define F(a, j) a[(long)(j) * nhor + h]
#define M(a, i, j) a[(long)(j) * n + (i)]
  for (long h = 0; h < nhor; ++h) {
    for (int j = 0; j < n; ++j) {
      zvgpg[j] = rcsq[h] * (F(gu, j) * gpm[h] + F(gv, j) * gpj[h]);
      double zq;
      if (nqspec == 1) {
        zq = F(gq, j) / gp[h];
        if (!(zq > 0.0)) zq = 0.0;
      } else {
        zq = F(dq, j);
        if (!(zq > 0.0)) zq = 0.0;
      }
      ztv1[j] = (F(gt, j) + t0[j]) * (1.0 + (1.0 / rdbrv - 1.0) * zq) - t0[j];
      ztv2[j] = (F(gt, j) + t0[j]) * (1.0 + (1.0 / rdbrv - 1.0) * zq) / (1.0 + adv * zq) - t0[j];
    }
    double zsumd = dsigma[0] * F(gd, 0);
    double vp = dsigma[0] * zvgpg[0];
    zsdotp[0] = zsumd + vp;
    for (int j = 1; j < n - 1; ++j) {
      zsumd = zsumd + dsigma[j] * F(gd, j);
      vp = vp + dsigma[j] * zvgpg[j];
      zsdotp[j] = zsumd + vp;
    }
    zsumd = zsumd + dsigma[n - 1] * F(gd, n - 1);
    vp = vp + dsigma[n - 1] * zvgpg[n - 1];
    gvp[h] = vp;
    for (int j = 0; j < n - 1; ++j) {
      zsdotp[j] = sigmah[j] * (zsumd + vp) - zsdotp[j];
      gtd[j] = zsdotp[j] * (F(gt, j + 1) - F(gt, j));
      gqm[j] = zsdotp[j] * (F(gq, j + 1) + F(gq, j));
      gud[j] = zsdotp[j] * (F(gu, j + 1) - F(gu, j));
      gvd[j] = zsdotp[j] * (F(gv, j + 1) - F(gv, j));
    }
    /* top level */
    zsumd = zvgpg[0] * dsigma[0];
    F(gtn, 0) = F(gt, 0) * F(gd, 0) - akap * ztv2[0] * F(gd, 0) -
                rdsig[0] * (gtd[0] + t01s2[0] * (sigmah[0] * vp - zsumd));
    F(gqn, 0) = -rdsig[0] * gqm[0];
    F(guz, 0) = -F(gu, 0) * F(gz, 0) - gpj[h] * ztv1[0] - rdsig[0] * gvd[0];
    F(gvz, 0) = F(gv, 0) * F(gz, 0) - gpm[h] * ztv1[0] - rdsig[0] * gud[0];
    F(dw, 0) = F(gd, 0) * gp[h] * psurf * ww;
    /* inner levels */
    for (int j = 1; j < n - 1; ++j) {
      double ztpta = M(c, 0, j) * zvgpg[0];
      double ztptb = M(c, 0, j) * (zvgpg[0] + F(gd, 0));
      for (int k = 1; k <= j; ++k) {
        ztpta = ztpta + M(c, k, j) * zvgpg[k];
        ztptb = ztptb + M(c, k, j) * (zvgpg[k] + F(gd, k));
      }
      zsumd = zsumd + zvgpg[j] * dsigma[j];
      F(gtn, j) = F(gt, j) * F(gd, j) + akap * ztv2[j] * (zvgpg[j] - ztptb) +
                  tkp[j] * (zvgpg[j] - ztpta) -
                  rdsig[j] * (gtd[j] + gtd[j - 1] +
                              vp * (t01s2[j] * sigmah[j] + t01s2[j - 1] * sigmah[j - 1]) -
                              zsumd * (t01s2[j - 1] + t01s2[j]) +
                              zvgpg[j] * dsigma[j] * t01s2[j - 1]);
      F(gqn, j) = -rdsig[j] * (gqm[j] - gqm[j - 1]);
      F(guz, j) = -F(gu, j) * F(gz, j) - gpj[h] * ztv1[j] - rdsig[j] * (gvd[j] + gvd[j - 1]);
      F(gvz, j) = F(gv, j) * F(gz, j) - gpm[h] * ztv1[j] - rdsig[j] * (gud[j] + gud[j - 1]);
      F(dw, j) = (zvgpg[j] - ztptb) * gp[h] * psurf * ww;
    }
    /* bottom level */
    {
      const int j = n - 1;
      double ztpta = M(c, 0, j) * zvgpg[0];
      double ztptb = M(c, 0, j) * (zvgpg[0] + F(gd, 0));
      for (int k = 1; k <= j; ++k) {
        ztpta = ztpta + M(c, k, j) * zvgpg[k];
        ztptb = ztptb + M(c, k, j) * (zvgpg[k] + F(gd, k));
      }
      F(gtn, j) = F(gt, j) * F(gd, j) + akap * ztv2[j] * (zvgpg[j] - ztptb) +
                  tkp[j] * (zvgpg[j] - ztpta) -
                  rdsig[j] * (gtd[j - 1] + t01s2[j - 1] * (sigmah[j - 1] * vp - zsumd));
      F(gqn, j) = rdsig[j] * gqm[j - 1];
      F(guz, j) = -F(gu, j) * F(gz, j) - gpj[h] * ztv1[j] - rdsig[j] * gvd[j - 1];
      F(gvz, j) = F(gv, j) * F(gz, j) - gpm[h] * ztv1[j] - rdsig[j] * gud[j - 1];
      F(dw, j) = (zvgpg[j] - ztptb) * gp[h] * psurf * ww;
    }
    /* nonlinear geopotential terms */
    for (int j = 0; j < n; ++j) ztv1[j] = ztv1[j] - F(gt, j);
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s = s + M(g, k, j) * ztv1[k];
      F(gphi, j) = s * 2.0 / rcsq[h];
    }
  }
#undef F
#undef M
  free(zv);
}
