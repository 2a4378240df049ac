// rad_lwr.hip -- the long-wave radiation column of the reference
// (src/radmod.f90:2122-2456, subroutine lwr) and its heating rate
// (src/radmod.f90:1100-1108) as a GPU kernel: one column per group of lanes,
// one lane per level.
//
// Arrays use the reference layout field(NHOR, lev): the column index is
// fastest.  The column work is O(NLEV^2) pow/log evaluations (absorber paths
// from every level to every level below), so this kernel is bound by the
// transcendental throughput of the vector ALUs.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/exo_rad.h"

namespace {

constexpr int MAXLEV = 64;

template <typename R>
struct LwrArgs {
  int nlev, ncol;
  const double *sigma, *sigmah, *dsigma;
  double ga, th2oc, tpofmt, acllwr, clgray, acpd, adv;
  int npbroaden;
  const R *dp, *dt, *dq, *dqo3, *dqco2, *dcc, *dql, *dls;
  R *dftu, *dftd, *heat, *olr;  // heat in K/day; any output may be null
};

template <typename R>
__device__ __forceinline__ R pw(R x, R y) {
  if constexpr (std::is_same<R, double>::value) return pow(x, y); else return powf(x, y);
}
template <typename R>
__device__ __forceinline__ R lg10(R x) {
  if constexpr (std::is_same<R, double>::value) return log10(x); else return log10f(x);
}
template <typename R>
__device__ __forceinline__ R ln(R x) {
  if constexpr (std::is_same<R, double>::value) return log(x); else return logf(x);
}
template <typename R>
__device__ __forceinline__ R ex(R x) {
  if constexpr (std::is_same<R, double>::value) return exp(x); else return expf(x);
}

// One column per group of LPC lanes (LPC = 32 for NLEV <= 31, else 64) and
// CPB columns per block, lane
// j - 1 owning level j: the absorber paths from level jlev downwards are a
// sequential sum per jlev, so lane jlev runs its own j2 loop and leaves its
// row of transmissivities ztau(jlev, j2) in LDS; the flux sums then read the
// rows (dftu, lane jlev) or the columns (dftd, lane j) of that table.  Every
// accumulated value sees its terms in the reference's order.
template <typename R>
struct LwrLds {
  // per column: zst4, zst4h, zbd, zbu, dftu, dftd [NLEV+2]; zq, zqo3, zqco2,
  // ztaucc0, ztaucs (row jlev = 1), ztausf [NLEV+1]; ztau [NLEV][NLEV]
  static __host__ __device__ int reals(int nlev) { return 12 * (nlev + 2) + nlev * nlev + 1; }
};

template <typename R, int LPC, int CPB>
__global__ __launch_bounds__(256) void k_lwr(LwrArgs<R> a) {
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x % LPC, cb = threadIdx.x / LPC;
  const int c = blockIdx.x * CPB + cb;
  const bool live = c < a.ncol;  // no early return: the block meets at barriers
  const int NLEV = a.nlev, NLEP = a.nlev + 1, NLEM = a.nlev - 1, NC = a.ncol;
  const int L2 = NLEV + 2, L1 = NLEV + 1;
  R* base = (R*)smem + (long)cb * LwrLds<R>::reals(NLEV);
  R *zst4 = base, *zst4h = zst4 + L2, *zbd = zst4h + L2, *zbu = zbd + L2;
  R *dftu = zbu + L2, *dftd = dftu + L2;
  R *zq = dftd + L2, *zqo3 = zq + L1, *zqco2 = zqo3 + L1, *ztaucc0 = zqco2 + L1;
  R *ztaucs = ztaucc0 + L1, *ztausf = ztaucs + L1, *ztau = ztausf + L1;  // ztau[(jlev-1)*NLEV + j2-1]
  R* zeps2 = ztau + NLEV * NLEV;
  auto TAU = [&](int jlev, int j2) -> R& { return ztau[(jlev - 1) * NLEV + (j2 - 1)]; };
  const double* S = a.sigma - 1;
  const double* SH = a.sigmah - 1;
  const double* DS = a.dsigma - 1;
  const R SBK = R(5.67e-8);
  const R zero = R(1.e-6);
  const R zpv2pm = R(0.0440098 / 0.0289644);
  const R zfh2o = R(0.1), zfco2 = R(100. / 1.9635), zfo3 = R(100. / 2.14);
  const int j = lane + 1;  // this lane's level
  const R ps = live ? a.dp[c] : R(1);
  const R ls = live ? a.dls[c] : R(0);
  const R zeps = ls + R(0.98) * (R(1.) - ls);

  // ---- level arrays ----
  if (live && j <= NLEP) {
    const R t = a.dt[(long)(j - 1) * NC + c];
    zst4[j] = SBK * (t * t) * (t * t);
  }
  __syncthreads();
  if (live && j <= NLEV) {
    if (j == 1) {
      zst4h[1] = zst4[1] - (zst4[1] - zst4[2]) * R(S[1]) / R(S[1] - S[2]);
    } else {
      const int jm = j - 1;
      zst4h[j] = (zst4[j] * R(S[jm] - SH[jm]) + zst4[jm] * R(SH[jm] - S[j])) / R(S[jm] - S[j]);
    }
    const R zps2 = (a.npbroaden > 0.5) ? ps * ps : R(101100.0) * R(101100.0);
    const long o = (long)(j - 1) * NC + c;
    const R zzf1 = R(S[j] * DS[j] / a.ga / 100000.);
    const R zzf3 = R(-1.66 * a.acllwr * 1000. * DS[j] / a.ga);
    const R zsfac = zzf1 * zps2;
    zq[j] = zfh2o * zsfac * a.dq[o];
    zqo3[j] = zfo3 * zsfac * a.dqo3[o];
    zqco2[j] = zfco2 * (zpv2pm * R(1.e-6)) * zsfac * a.dqco2[o];
    if (a.clgray > 0)
      ztaucc0[j] = R(1.) - a.dcc[o] * R(a.clgray);
    else
      ztaucc0[j] = R(1.) - a.dcc[o] * (R(1.) - ex(zzf3 * a.dql[o] * ps));
  }
  __syncthreads();

  // ---- transmissivities from level jlev = j downwards ----
  if (live && j <= NLEV) {
    const R zao30 = R(0.209) * pw(R(7.e-5), R(0.436));
    const R zco20 = R(0.0676) * pw(R(0.01022), R(0.421));
    const R zh2o0a = R(0.846) * pw(R(3.59e-5), R(0.243));
    const R zh2o0 = R(0.832) * pw(R(0.0286), R(0.26));
    const R zao3c = R(0.209) * pw(R(0.01 + 7.e-5), R(0.436)) - zao30 - R(0.0212) * lg10(R(0.01));
    const R zaco2c = R(0.0676) * pw(R(1.01022), R(0.421)) - zco20;
    const R zah2oc =
        R(0.846) * pw(R(0.01 + 3.59e-5), R(0.243)) - zh2o0a - R(0.24) * lg10(R(0.02));
    const R zth2oc = R(1.) - (R(0.832) * pw(R(2. + 0.0286), R(0.26)) - zh2o0) +
                     R(0.1196) * ln(R(2. - 0.6931));
    const int jlev = j;
    R ztaucc = 1, zsumwv = 0, zsumo3 = 0, zsumco2 = 0;
    for (int j2 = jlev; j2 <= NLEV; ++j2) {
      zsumwv += zq[j2];
      zsumo3 += zqo3[j2];
      zsumco2 += zqco2[j2];
      R zah2o, zaco2, zth2o, zao3;
      if (zsumwv <= R(0.01))
        zah2o = R(0.846) * pw(zsumwv + R(3.59e-5), R(0.243)) - zh2o0a;
      else
        zah2o = R(0.24) * lg10(zsumwv + R(0.01)) + zah2oc;
      if (a.th2oc > 0.) zah2o = fmin(zah2o + (R(1.) - ex(-R(a.th2oc) * zsumwv)), R(1.));
      if (zsumco2 <= R(1.0))
        zaco2 = R(0.0676) * pw(zsumco2 + R(0.01022), R(0.421)) - zco20;
      else
        zaco2 = R(0.0546) * lg10(zsumco2) + zaco2c;
      if (zsumwv <= R(2.))
        zth2o = R(1.) - (R(0.832) * pw(zsumwv + R(0.0286), R(0.26)) - zh2o0);
      else
        zth2o = fmax(R(0.), zth2oc - R(0.1196) * ln(zsumwv - R(0.6931)));
      if (zsumo3 <= R(0.01))
        zao3 = R(0.209) * pw(zsumo3 + R(7.e-5), R(0.436)) - zao30;
      else
        zao3 = R(0.0212) * lg10(zsumo3) + zao3c;
      R tcs = R(1.) - zah2o - zao3 - zaco2 * zth2o;
      tcs = fmin(R(1.) - zero, fmax(zero, tcs));
      if (jlev == 1) ztaucs[j2] = tcs;
      ztaucc = ztaucc * ztaucc0[j2];
      TAU(jlev, j2) = tcs * ztaucc;
    }
    ztausf[jlev] = TAU(jlev, NLEV);
  }
  if (live && lane == 0) {
    if ((zst4[NLEV] - zst4h[NLEV]) * (zst4[NLEV] - zst4[NLEP]) > 0)
      zst4h[NLEP] = zst4[NLEP];
    else
      zst4h[NLEP] =
          zst4[NLEV] + (zst4[NLEV] - zst4[NLEM]) * R(1. - S[NLEV]) / R(S[NLEV] - S[NLEM]);
    zbd[0] = 0;  // SBK * zttop**4 with zttop = 0
    zbu[NLEP] = zeps * zst4[NLEP];
  }
  __syncthreads();

  // ---- effective radiation of each layer (the jlev = 1 pass of the
  // reference).  The transmissivity carried into layer j2 is ztaucs of the
  // layer above, so the layers are independent.  Near ztau0 = 1 - 1e-6
  // both 1/(1 - ztau0) and 1/ln(ztau0) are ~1e6 and cancel; the function is
  // well conditioned, its evaluation is not, so this runs in FP64 for
  // either column precision. ----
  if (live && j <= NLEV) {
    const double zd = 1.e-6;
    const int j2 = j, jp2 = j2 + 1;
    double ztau0 = (j2 == 1) ? 1. - zd : (double)ztaucs[j2 - 1];
    ztau0 = (double)ztaucs[j2] / ztau0 * (double)ztaucc0[j2] * a.tpofmt;
    ztau0 = fmin(1. - zd, fmax(zd, ztau0));
    if ((zst4[j2] - zst4h[j2]) * (zst4[j2] - zst4h[jp2]) > 0) {
      zbd[j2] = R(0.5) * zst4[j2] + R(0.25) * (zst4h[j2] + zst4h[jp2]);
      zbu[j2] = zbd[j2];
    } else {
      const double h1 = zst4h[j2], h2 = zst4h[jp2];
      const double bd = (h2 - ztau0 * h1) / (1. - ztau0) - (h1 - h2) / log(ztau0);
      zbd[j2] = (R)bd;
      zbu[j2] = (R)(h1 + h2 - bd);
    }
  }
  __syncthreads();

  // ---- fluxes: dftu(jlev) over its row of ztau, dftd(j) over column j - 1 ----
  if (live && j <= NLEP) {
    if (j <= NLEV) {
      R du = R(0) - zbu[j];
      for (int j2 = j; j2 <= NLEV; ++j2) du = du - (zbu[j2 + 1] - zbu[j2]) * TAU(j, j2);
      dftu[j] = du;
    } else {
      dftu[NLEP] = R(0) - zbu[NLEP];
    }
    R dd = 0;
    for (int jlev = 1; jlev < j; ++jlev) dd = dd + (zbd[jlev - 1] - zbd[jlev]) * TAU(jlev, j - 1);
    dd = dd + zbd[j - 1];
    dftd[j] = dd;
    if (j == NLEP) *zeps2 = (R(1.) - zeps) * dd;
  }
  __syncthreads();
  if (live && j <= NLEV) dftu[j] = dftu[j] - ztausf[j] * (*zeps2);
  __syncthreads();
  if (!live || j > NLEP) return;
  if (a.dftu) a.dftu[(long)(j - 1) * NC + c] = dftu[j];
  if (a.dftd) a.dftd[(long)(j - 1) * NC + c] = dftd[j];
  if (a.heat && j <= NLEV) {
    const R f1 = dftu[j + 1] + dftd[j + 1], f0 = dftu[j] + dftd[j];
    const R q = a.dq[(long)(j - 1) * NC + c];
    a.heat[(long)(j - 1) * NC + c] =
        R(-86400.0 * a.ga) * (f1 - f0) / (R(DS[j]) * ps * R(a.acpd) * (R(1.) + R(a.adv) * q));
  }
  if (a.olr && j == 1) a.olr[c] = -dftu[1];
}

template <typename R>
int launch(const exo_lwr_params* p, int ncol, const exo_lwr_columns* in, exo_lwr_out* out,
           hipStream_t s) {
  LwrArgs<R> a{};
  a.nlev = p->nlev;
  a.ncol = ncol;
  a.sigma = p->sigma;
  a.sigmah = p->sigmah;
  a.dsigma = p->dsigma;
  a.ga = p->ga;
  a.th2oc = p->th2oc;
  a.tpofmt = p->tpofmt;
  a.acllwr = p->acllwr;
  a.clgray = p->clgray;
  a.acpd = p->acpd;
  a.adv = p->adv;
  a.npbroaden = p->npbroaden;
  a.dp = (const R*)in->dp;
  a.dt = (const R*)in->dt;
  a.dq = (const R*)in->dq;
  a.dqo3 = (const R*)in->dqo3;
  a.dqco2 = (const R*)in->dqco2;
  a.dcc = (const R*)in->dcc;
  a.dql = (const R*)in->dql;
  a.dls = (const R*)in->dls;
  a.dftu = (R*)out->dftu;
  a.dftd = (R*)out->dftd;
  a.heat = (R*)out->heat;
  a.olr = (R*)out->olr;
  const size_t per_col = sizeof(R) * LwrLds<R>::reals(p->nlev);
  // 8 columns of 32 lanes per block; above 31 levels one column of 64
  // lanes per block (its table of NLEV^2 transmissivities is up to 32 KB)
  if (p->nlev + 1 <= 32)
    k_lwr<R, 32, 8><<<(ncol + 7) / 8, 256, 8 * per_col, s>>>(a);
  else
    k_lwr<R, 64, 1><<<ncol, 64, per_col, s>>>(a);
  return hipGetLastError() == hipSuccess ? EXO_RAD_OK : EXO_RAD_ERR_HIP;
}

}  // namespace

extern "C" int exo_lwr(const exo_lwr_params* p, int fp64, int ncol, const exo_lwr_columns* in,
                       exo_lwr_out* out, void* stream) {
  if (!p || !in || !out || ncol < 0 || p->nlev < 2 || p->nlev > MAXLEV - 1 || !p->sigma ||
      !p->sigmah || !p->dsigma)
    return EXO_RAD_ERR_ARG;
  if (ncol == 0) return EXO_RAD_OK;
  if (!in->dp || !in->dt || !in->dq || !in->dqo3 || !in->dqco2 || !in->dcc || !in->dql ||
      !in->dls)
    return EXO_RAD_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  return fp64 ? launch<double>(p, ncol, in, out, s) : launch<float>(p, ncol, in, out, s);
}
