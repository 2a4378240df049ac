/*
 * exo_oracle_lwr.c -- FP64 CPU restatement of the long-wave radiation column
 * of the reference (src/radmod.f90:2122-2456, subroutine lwr) and of its
 * heating-rate conversion (src/radmod.f90:1100-1108).
 *
 * TEST INFRASTRUCTURE ONLY (the CPU column that the GPU column kernel and the
 * learned emulator are tested against); the product never links it.
 *
 * One call transforms ncol independent columns.  Arrays are [col][lev]
 * (one column contiguous); level 1 of the reference is index 0 (top) and
 * NLEP = NLEV+1 is the surface.  Pinned by tests/test_oracle_lwr.py against
 * the reference's lwr compiled from its sources (oracle/_ref) and against
 * the committed golden vectors written from it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "exo_oracle.h"

#define SBK 5.67e-8

void ora_lwr(int nlev, int ncol, const double *sigma, const double *sigmah,
             const double *dsigma, double ga, double th2oc, double tpofmt,
             double acllwr, double clgray, int npbroaden, const double *dp,
             const double *dt, const double *dq, const double *dqo3,
             const double *dqco2, const double *dcc, const double *dql,
             const double *dls, double *dftu_out, double *dftd_out) {
  const int NLEV = nlev, NLEP = nlev + 1, NLEM = nlev - 1;
  /* 1-based level arrays with room for index 0 */
  double *zbu = malloc(sizeof(double) * (NLEP + 1));
  double *zbd = malloc(sizeof(double) * (NLEP + 1));
  double *zst4h = malloc(sizeof(double) * (NLEP + 1));
  double *zst4 = malloc(sizeof(double) * (NLEP + 1));
  double *ztau = malloc(sizeof(double) * (NLEP + 1));
  double *zq = malloc(sizeof(double) * (NLEP + 1));
  double *zqo3 = malloc(sizeof(double) * (NLEP + 1));
  double *zqco2 = malloc(sizeof(double) * (NLEP + 1));
  double *ztausf = malloc(sizeof(double) * (NLEP + 1));
  double *ztaucs = malloc(sizeof(double) * (NLEP + 1));
  double *ztaucc0 = malloc(sizeof(double) * (NLEP + 1));
  double *dftu = malloc(sizeof(double) * (NLEP + 1));
  double *dftd = malloc(sizeof(double) * (NLEP + 1));
  const double *S = sigma - 1, *SH = sigmah - 1, *DS = dsigma - 1;

  /* local parameters (src/radmod.f90:2150-2168) */
  const double zmmair = 0.0289644, zmmco2 = 0.0440098;
  const double zpv2pm = zmmco2 / zmmair;
  const double zrco2 = 1.9635, zro3 = 2.14, zttop = 0.0;
  const double zfh2o = 0.1, zfco2 = 100. / zrco2, zfo3 = 100. / zro3;
  /* section 1 (src/radmod.f90:2200-2228) */
  const double zero = 1.e-6;
  const double zao30 = 0.209 * pow(7.e-5, 0.436);
  const double zco20 = 0.0676 * pow(0.01022, 0.421);
  const double zh2o0a = 0.846 * pow(3.59e-5, 0.243);
  const double zh2o0 = 0.832 * pow(0.0286, 0.26);
  const double zao3c = 0.209 * pow(0.01 + 7.e-5, 0.436) - zao30 - 0.0212 * log10(0.01);
  const double zaco2c = 0.0676 * pow(1.01022, 0.421) - zco20;
  const double zah2oc = 0.846 * pow(0.01 + 3.59e-5, 0.243) - zh2o0a - 0.24 * log10(0.02);
  const double zth2oc = 1. - (0.832 * pow(2. + 0.0286, 0.26) - zh2o0) + 0.1196 * log(2. - 0.6931);

  for (int c = 0; c < ncol; ++c) {
    const double *T = dt + (long)c * NLEP - 1; /* T[1..NLEP] */
    const double *Q = dq + (long)c * NLEV - 1;
    const double *O3 = dqo3 + (long)c * NLEV - 1;
    const double *CO2 = dqco2 + (long)c * NLEV - 1;
    const double *CC = dcc + (long)c * NLEV - 1;
    const double *QL = dql + (long)c * NLEV - 1;
    const double ps = dp[c];
    const double zps2 = (npbroaden > 0.5) ? ps * ps : 101100.0 * 101100.0;
    for (int j = 1; j <= NLEP; ++j) {
      dftu[j] = 0.;
      dftd[j] = 0.;
    }
    /* section 2 (src/radmod.f90:2240-2271) */
    for (int j = 1; j <= NLEP; ++j) zst4[j] = SBK * pow(T[j], 4);
    zst4h[1] = zst4[1] - (zst4[1] - zst4[2]) * S[1] / (S[1] - S[2]);
    for (int j = 2; j <= NLEV; ++j) {
      const int jm = j - 1;
      zst4h[j] = (zst4[j] * (S[jm] - SH[jm]) + zst4[jm] * (SH[jm] - S[j])) / (S[jm] - S[j]);
    }
    if ((zst4[NLEV] - zst4h[NLEV]) * (zst4[NLEV] - zst4[NLEP]) > 0.)
      zst4h[NLEP] = zst4[NLEP];
    else
      zst4h[NLEP] = zst4[NLEV] + (zst4[NLEV] - zst4[NLEM]) * (1. - S[NLEV]) / (S[NLEV] - S[NLEM]);
    zbd[0] = SBK * pow(zttop, 4);
    double zeps = dls[c] + 0.98 * (1. - dls[c]);
    zbu[NLEP] = zeps * zst4[NLEP];
    /* section 3a (src/radmod.f90:2281-2296) */
    for (int j = 1; j <= NLEV; ++j) {
      const double zzf1 = S[j] * DS[j] / ga / 100000.;
      const double zzf2 = zpv2pm * 1.e-6;
      const double zzf3 = -1.66 * acllwr * 1000. * DS[j] / ga;
      const double zsfac = zzf1 * zps2;
      zq[j] = zfh2o * zsfac * Q[j];
      zqo3[j] = zfo3 * zsfac * O3[j];
      zqco2[j] = zfco2 * zzf2 * zsfac * CO2[j];
      if (clgray > 0)
        ztaucc0[j] = 1. - CC[j] * clgray;
      else
        ztaucc0[j] = 1. - CC[j] * (1. - exp(zzf3 * QL[j] * ps));
    }
    /* section 3b (src/radmod.f90:2298-2430) */
    for (int jlev = 1; jlev <= NLEV; ++jlev) {
      const int jlem = jlev - 1;
      double ztaucc = 1., zsumwv = 0., zsumo3 = 0., zsumco2 = 0.;
      for (int j2 = jlev; j2 <= NLEV; ++j2) {
        zsumwv += zq[j2];
        zsumo3 += zqo3[j2];
        zsumco2 += zqco2[j2];
        double zah2o, zaco2, zth2o, zao3;
        if (zsumwv <= 0.01)
          zah2o = 0.846 * pow(zsumwv + 3.59e-5, 0.243) - zh2o0a;
        else
          zah2o = 0.24 * log10(zsumwv + 0.01) + zah2oc;
        if (th2oc > 0.) zah2o = fmin(zah2o + (1. - exp(-th2oc * zsumwv)), 1.);
        if (zsumco2 <= 1.0)
          zaco2 = 0.0676 * pow(zsumco2 + 0.01022, 0.421) - zco20;
        else
          zaco2 = 0.0546 * log10(zsumco2) + zaco2c;
        if (zsumwv <= 2.)
          zth2o = 1. - (0.832 * pow(zsumwv + 0.0286, 0.26) - zh2o0);
        else
          zth2o = fmax(0., zth2oc - 0.1196 * log(zsumwv - 0.6931));
        if (zsumo3 <= 0.01)
          zao3 = 0.209 * pow(zsumo3 + 7.e-5, 0.436) - zao30;
        else
          zao3 = 0.0212 * log10(zsumo3) + zao3c;
        double tcs = 1. - zah2o - zao3 - zaco2 * zth2o;
        tcs = fmin(1. - zero, fmax(zero, tcs));
        ztaucs[j2] = tcs;
        ztaucc = ztaucc * ztaucc0[j2];
        ztau[j2] = tcs * ztaucc;
      }
      if (jlev == 1) {
        double ztau0 = 1. - zero;
        for (int j2 = 1; j2 <= NLEV; ++j2) {
          const int jp2 = j2 + 1;
          ztau0 = ztaucs[j2] / ztau0 * ztaucc0[j2] * tpofmt;
          ztau0 = fmin(1. - zero, fmax(zero, ztau0));
          if ((zst4[j2] - zst4h[j2]) * (zst4[j2] - zst4h[jp2]) > 0.) {
            zbd[j2] = 0.5 * zst4[j2] + 0.25 * (zst4h[j2] + zst4h[jp2]);
            zbu[j2] = zbd[j2];
          } else {
            zbd[j2] = (zst4h[jp2] - ztau0 * zst4h[j2]) / (1. - ztau0) -
                      (zst4h[j2] - zst4h[jp2]) / log(ztau0);
            zbu[j2] = zst4h[j2] + zst4h[jp2] - zbd[j2];
          }
          ztau0 = ztaucs[j2];
        }
      }
      dftu[jlev] = dftu[jlev] - zbu[jlev];
      dftd[jlev] = dftd[jlev] + zbd[jlem];
      const double zbdl = zbd[jlem] - zbd[jlev];
      for (int j2 = jlev; j2 <= NLEV; ++j2) {
        const int jp2 = j2 + 1;
        dftu[jlev] = dftu[jlev] - (zbu[jp2] - zbu[j2]) * ztau[j2];
        dftd[jp2] = dftd[jp2] + zbdl * ztau[j2];
      }
      ztausf[jlev] = ztau[NLEV];
    }
    /* surface and non-black surface (src/radmod.f90:2432-2449) */
    dftu[NLEP] = dftu[NLEP] - zbu[NLEP];
    dftd[NLEP] = dftd[NLEP] + zbd[NLEV];
    zeps = (1. - zeps) * dftd[NLEP];
    for (int j = 1; j <= NLEV; ++j) dftu[j] = dftu[j] - ztausf[j] * zeps;
    for (int j = 1; j <= NLEP; ++j) {
      dftu_out[(long)c * NLEP + j - 1] = dftu[j];
      dftd_out[(long)c * NLEP + j - 1] = dftd[j];
    }
  }
  free(zbu); free(zbd); free(zst4h); free(zst4); free(ztau); free(zq); free(zqo3);
  free(zqco2); free(ztausf); free(ztaucs); free(ztaucc0); free(dftu); free(dftd);
}

/* heating rate dtdtlwr = -ga dF / (dsigma ps acpd (1 + adv q)) in K/s
 * (src/radmod.f90:1107-1108) from the net flux dlwfl = dftu + dftd */
void ora_lwr_heating(int nlev, int ncol, const double *dsigma, double ga, double acpd,
                     double adv, const double *dp, const double *dq, const double *dftu,
                     const double *dftd, double *dtdt) {
  const int NLEP = nlev + 1;
  for (int c = 0; c < ncol; ++c) {
    const double *u = dftu + (long)c * NLEP, *d = dftd + (long)c * NLEP;
    for (int j = 0; j < nlev; ++j) {
      const double f1 = u[j + 1] + d[j + 1], f0 = u[j] + d[j];
      dtdt[(long)c * nlev + j] =
          -ga * (f1 - f0) / (dsigma[j] * dp[c] * acpd * (1. + adv * dq[(long)c * nlev + j]));
    }
  }
}

/* sigma levels of the neqsig = 0 model configuration
 * (src/plasim.f90:1553-1567) */
void ora_sigma_levels(int nlev, double *sigma, double *sigmah, double *dsigma) {
  for (int j = 1; j <= nlev; ++j) {
    const double zsk = (double)j / (double)nlev;
    sigmah[j - 1] = 0.75 * zsk + 1.75 * (zsk * zsk * zsk) - 1.5 * ((zsk * zsk) * (zsk * zsk));
  }
  dsigma[0] = sigmah[0];
  for (int j = 1; j < nlev; ++j) dsigma[j] = sigmah[j] - sigmah[j - 1];
  sigma[0] = 0.5 * sigmah[0];
  for (int j = 1; j < nlev; ++j) sigma[j] = 0.5 * (sigmah[j - 1] + sigmah[j]);
}
