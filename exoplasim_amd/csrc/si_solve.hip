// si_solve.hip -- the semi-implicit spectral solve in FP64 on the GPU.
//
//   k_makebm      bm1(:,:,n) = inverse(delt^2 (t0 dsigma + g.tau) + I/(n(n+1)))
//                 for n = 1..NTRU (src/plasim.f90:1683-1703), one workgroup
//                 per n, Gauss-Jordan elimination with partial pivoting in LDS.
//   k_si_apply    divergence at time t from eq. 17 of Hoskins & Simmons
//                 (1975) and the ln(ps) and temperature updates
//                 (src/plasim.f90:2817-2876), one thread per spectral real.
// Matrices are column-major NLEV x NLEV as in the reference.
#include <hip/hip_runtime.h>

#include "../../include/exo_si.h"

namespace {

constexpr int MAXLEV = 64;

__global__ void k_makebm(int nlev, int ntru, double delt, const double* t0,
                         const double* dsigma, const double* g, const double* tau,
                         double* bm1, int* status) {
  const int jn = blockIdx.x + 1;  // total wavenumber 1..ntru
  if (jn > ntru) return;
  __shared__ double a[MAXLEV][2 * MAXLEV + 1];  // [row][col], augmented [A | I]
  __shared__ int piv;
  const int tid = threadIdx.x;
  const int n2 = 2 * nlev;
  const double zdeltsq = delt * delt;
  // A(j2, j1) = delt^2 (t0(j1) dsigma(j2) + g(:,j1).tau(j2,:)) + delta/(n(n+1))
  for (int idx = tid; idx < nlev * nlev; idx += blockDim.x) {
    const int j2 = idx % nlev, j1 = idx / nlev;
    double dp = 0.0;
    for (int k = 0; k < nlev; ++k) dp += g[j1 * nlev + k] * tau[k * nlev + j2];
    double v = zdeltsq * (t0[j1] * dsigma[j2] + dp);
    if (j1 == j2) v += 1.0 / (double)(jn * (jn + 1));
    a[j2][j1] = v;
    a[j2][nlev + j1] = (j1 == j2) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int col = 0; col < nlev; ++col) {
    if (tid == 0) {
      int p = col;
      double best = fabs(a[col][col]);
      for (int r = col + 1; r < nlev; ++r)
        if (fabs(a[r][col]) > best) {
          best = fabs(a[r][col]);
          p = r;
        }
      piv = p;
      if (best == 0.0) atomicExch(status, 1);
    }
    __syncthreads();
    const int p = piv;
    if (p != col)
      for (int c = tid; c < n2; c += blockDim.x) {
        const double t = a[col][c];
        a[col][c] = a[p][c];
        a[p][c] = t;
      }
    __syncthreads();
    const double inv = 1.0 / a[col][col];
    __syncthreads();
    for (int c = tid; c < n2; c += blockDim.x) a[col][c] *= inv;
    __syncthreads();
    for (int idx = tid; idx < nlev * n2; idx += blockDim.x) {
      const int r = idx / n2, c = idx % n2;
      if (r != col && c != col) a[r][c] -= a[r][col] * a[col][c];
    }
    __syncthreads();
    for (int r = tid; r < nlev; r += blockDim.x)
      if (r != col) a[r][col] = 0.0;
    __syncthreads();
  }
  double* out = bm1 + (long)(jn - 1) * nlev * nlev;
  for (int idx = tid; idx < nlev * nlev; idx += blockDim.x) {
    const int r = idx % nlev, c = idx / nlev;
    out[c * nlev + r] = a[r][nlev + c];
  }
}

// interface arrays of type R (the model's real kind); arithmetic in FP64
template <typename R>
__global__ void k_si_apply(int nlev, int nsp, const int* nindex, double delt, const double* t0,
                           const double* dsigma, const double* g, const double* tau,
                           const double* bm1, const R* sop, const R* spm, const R* atm,
                           const R* adm, R* stt, R* sdt, R* spt) {
  const int jsp = blockIdx.x * blockDim.x + threadIdx.x;
  if (jsp >= nsp) return;
  const int jn = nindex[jsp];
  double st[MAXLEV], zz[MAXLEV], d[MAXLEV];
  for (int l = 0; l < nlev; ++l) st[l] = (double)stt[(long)l * nsp + jsp];
  const double zq = jn > 0 ? 1.0 / (double)(jn * (jn + 1)) : 0.0;
  const double sp_t = spt[jsp], sp_m = spm[jsp], so = sop[jsp];
  // rhs of eq. 17 per level jlev2 (the same for every output level)
  for (int l2 = 0; l2 < nlev; ++l2) {
    double zgt = 0, zgm = 0;
    for (int k = 0; k < nlev; ++k) {
      zgt += g[l2 * nlev + k] * st[k];
      zgm += g[l2 * nlev + k] * (double)atm[(long)k * nsp + jsp];
    }
    const double z0 = t0[l2];
    const double zt = zgt - z0 * sp_t;
    const double zm = zgm + z0 * sp_m;
    const double za = (double)sdt[(long)l2 * nsp + jsp] * zq + so;
    const double zb = (double)adm[(long)l2 * nsp + jsp] * zq;
    zz[l2] = zb + delt * (zm + za + delt * zt);
  }
  for (int l = 0; l < nlev; ++l) {
    double zsum = 0;
    if (jn > 0) {
      const double* b = bm1 + (long)(jn - 1) * nlev * nlev + (long)l * nlev;
      for (int l2 = 0; l2 < nlev; ++l2) zsum += zz[l2] * b[l2];
    }
    d[l] = (jsp < 2) ? 0.0 : zsum;  // first mode is zero (src/plasim.f90:2860)
  }
  double ps = sp_t;
  for (int l = 0; l < nlev; ++l) ps += dsigma[l] * d[l];
  for (int l = 0; l < nlev; ++l) {
    double s = 0;
    for (int k = 0; k < nlev; ++k) s += tau[l * nlev + k] * d[k];
    stt[(long)l * nsp + jsp] = (R)(st[l] - s);
    sdt[(long)l * nsp + jsp] = (R)d[l];
  }
  spt[jsp] = (R)ps;
}

}  // namespace

struct exo_si_s {
  int nlev = 0, ntru = 0, nsp = 0;
  double *t0 = nullptr, *dsigma = nullptr, *g = nullptr, *tau = nullptr, *bm1 = nullptr;
  int* nindex = nullptr;
  int* status = nullptr;
};

#define HIPCK(x)                              \
  do {                                        \
    if ((x) != hipSuccess) return EXO_SI_ERR_HIP; \
  } while (0)

extern "C" int exo_si_create(exo_si_t* si, int nlev, int ntru, const double* t0,
                             const double* dsigma, const double* g, const double* tau) {
  if (!si || nlev < 1 || nlev > MAXLEV || ntru < 1 || !t0 || !dsigma || !g || !tau)
    return EXO_SI_ERR_ARG;
  exo_si_s* s = new exo_si_s();
  s->nlev = nlev;
  s->ntru = ntru;
  s->nsp = (ntru + 1) * (ntru + 2);
  const size_t m2 = sizeof(double) * nlev * nlev;
  HIPCK(hipMalloc(&s->t0, sizeof(double) * nlev));
  HIPCK(hipMalloc(&s->dsigma, sizeof(double) * nlev));
  HIPCK(hipMalloc(&s->g, m2));
  HIPCK(hipMalloc(&s->tau, m2));
  HIPCK(hipMalloc(&s->bm1, m2 * ntru));
  HIPCK(hipMalloc(&s->nindex, sizeof(int) * s->nsp));
  HIPCK(hipMalloc(&s->status, sizeof(int)));
  HIPCK(hipMemcpy(s->t0, t0, sizeof(double) * nlev, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(s->dsigma, dsigma, sizeof(double) * nlev, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(s->g, g, m2, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(s->tau, tau, m2, hipMemcpyHostToDevice));
  // total wavenumber of each spectral real (src/plasim.f90:1640-1646)
  int* ni = new int[s->nsp];
  for (int m = 0, jr = 0; m <= ntru; ++m)
    for (int n = m; n <= ntru; ++n, jr += 2) ni[jr] = ni[jr + 1] = n;
  hipError_t e = hipMemcpy(s->nindex, ni, sizeof(int) * s->nsp, hipMemcpyHostToDevice);
  delete[] ni;
  if (e != hipSuccess) return EXO_SI_ERR_HIP;
  *si = s;
  return EXO_SI_OK;
}

extern "C" int exo_si_destroy(exo_si_t s) {
  if (!s) return EXO_SI_ERR_ARG;
  void* ps[] = {s->t0, s->dsigma, s->g, s->tau, s->bm1, s->nindex, s->status};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete s;
  return EXO_SI_OK;
}

extern "C" int exo_si_makebm(exo_si_t s, double delt, void* stream) {
  if (!s) return EXO_SI_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  HIPCK(hipMemsetAsync(s->status, 0, sizeof(int), st));
  k_makebm<<<s->ntru, 64, 0, st>>>(s->nlev, s->ntru, delt, s->t0, s->dsigma, s->g, s->tau,
                                   s->bm1, s->status);
  HIPCK(hipGetLastError());
  int h = 0;
  HIPCK(hipMemcpyAsync(&h, s->status, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCK(hipStreamSynchronize(st));
  return h ? EXO_SI_ERR_SINGULAR : EXO_SI_OK;
}

extern "C" int exo_si_bm1(exo_si_t s, double* bm1) {
  if (!s || !bm1) return EXO_SI_ERR_ARG;
  HIPCK(hipMemcpy(bm1, s->bm1, sizeof(double) * s->nlev * s->nlev * s->ntru,
                  hipMemcpyDeviceToHost));
  return EXO_SI_OK;
}

template <typename R>
static int apply_t(exo_si_t s, double delt, const R* sop, const R* spm, const R* atm,
                   const R* adm, R* stt, R* sdt, R* spt, hipStream_t st) {
  k_si_apply<R><<<(s->nsp + 127) / 128, 128, 0, st>>>(s->nlev, s->nsp, s->nindex, delt,
                                                      s->t0, s->dsigma, s->g, s->tau, s->bm1,
                                                      sop, spm, atm, adm, stt, sdt, spt);
  return hipGetLastError() == hipSuccess ? EXO_SI_OK : EXO_SI_ERR_HIP;
}

extern "C" int exo_si_apply(exo_si_t s, int fp64, double delt, const void* sop, const void* spm,
                            const void* atm, const void* adm, void* stt, void* sdt, void* spt,
                            void* stream) {
  if (!s || !sop || !spm || !atm || !adm || !stt || !sdt || !spt) return EXO_SI_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (fp64)
    return apply_t<double>(s, delt, (const double*)sop, (const double*)spm, (const double*)atm,
                           (const double*)adm, (double*)stt, (double*)sdt, (double*)spt, st);
  return apply_t<float>(s, delt, (const float*)sop, (const float*)spm, (const float*)atm,
                        (const float*)adm, (float*)stt, (float*)sdt, (float*)spt, st);
}
