// sht_direct_terms.hip -- folded direct Legendre transforms with several
// Fourier inputs and tables (src/legmod.f90 folded branches):
//   uv2dv  (:391-422)   (u, v)                       -> (D, zeta)
//   qtend  (:468-496)   (qn, uq, vq)                 -> q
//   mktend (:555-602)   (tn, fu, fv, ke, ut, vt)     -> (D, T, zeta)
// Each output field is a sum of terms  sign * t[pair][w] * X_i[m][cls][pair][c]
// where X_i is input i folded into E = N+S (class sym) / O = N-S (class anti)
// and the class of table t at (m,n) is sym iff (n-m even) != (t is Q-type).
// One launch computes one output field (re and im) for every level; the
// caller launches once per output.
#include <type_traits>

#include "sht_internal.h"

namespace exo {

template <int MODE>
using RealX = typename std::conditional<MODE == MODE_F64, double, float>::type;

// latitude pairs per block so that 3 folded inputs fit in about 100 KB of LDS
template <int T, int MODE>
struct DirxCfg {
  static constexpr int PB = (T == 21) ? 16
                          : (T == 42) ? (MODE == MODE_F64 ? 16 : 32)
                          : (T == 63) ? (MODE == MODE_F64 ? 12 : 24)
                                      : (MODE == MODE_F64 ? 8 : 16);
  static constexpr int NPB = TS<T>::NPAIR / PB;
};

template <int T, int MODE>
__global__ __launch_bounds__(256) void k_dirx(DirxArgs a, const void* tabv, void* scratch) {
  using S = TS<T>;
  using R = RealX<MODE>;
  constexpr int NM = S::NM, N = S::NLON, NLAT = S::NLAT;
  constexpr int PB = DirxCfg<T, MODE>::PB, NPB = DirxCfg<T, MODE>::NPB;
  // X[i][m][cls][pair][ri]
  __shared__ R X[3 * NM * 2 * PB * 2];
  __shared__ uint8_t mw[S::NCSP];
  const int tid = threadIdx.x;
  const int pb = blockIdx.x % NPB;
  const int ml = blockIdx.x / NPB;  // member * nlev + level
  if (ml >= a.nmem * a.nlev) return;
  const int mem = ml / a.nlev, lev = ml - mem * a.nlev;
  for (int m = tid; m <= T; m += 256)
    for (int n = m; n <= T; ++n) mw[w0_of(T, m) + n - m] = (uint8_t)m;
  auto xi = [](int i, int m, int cls, int pl, int ri) {
    return (((i * NM + m) * 2 + cls) * PB + pl) * 2 + ri;
  };
  for (int idx = tid; idx < a.nin * PB * NM; idx += 256) {
    const int pl = idx % PB, k = (idx / PB) % NM, i = idx / (PB * NM);
    const int lat = pb * PB + pl;
    const R* src = (const R*)a.in[i] + mem * a.in_stride + (long)lev * NLAT * N;
    const R nx = src[(long)lat * N + 2 * k], ny = src[(long)lat * N + 2 * k + 1];
    const R sx = src[(long)(NLAT - 1 - lat) * N + 2 * k];
    const R sy = src[(long)(NLAT - 1 - lat) * N + 2 * k + 1];
    X[xi(i, k, 0, pl, 0)] = nx + sx;
    X[xi(i, k, 0, pl, 1)] = ny + sy;
    X[xi(i, k, 1, pl, 0)] = nx - sx;
    X[xi(i, k, 1, pl, 1)] = ny - sy;
  }
  __syncthreads();
  const R* tv = (const R*)tabv;
  const long tsz = (long)S::NPAIR * S::NCSP;
  R* out;
  if (NPB == 1)
    out = (R*)a.out + mem * a.out_stride + (long)lev * S::NRSP;
  else
    out = (R*)scratch + ((long)pb * a.nmem * a.nlev + ml) * S::NRSP;
  for (int w = tid; w < S::NCSP; w += 256) {
    const int m = mw[w];
    const int n = m + (w - w0_of(T, m));
    const bool even = ((n - m) & 1) == 0;
    R acc[2] = {R(0), R(0)};
    for (int t = 0; t < a.nterm; ++t) {
      const DirxTerm tm = a.terms[t];
      const int cls = (even != table_is_q(tm.table)) ? 0 : 1;
      const R* tp = tv + tm.table * tsz + (long)(pb * PB) * S::NCSP + w;
      R s = R(0);
      for (int pl = 0; pl < PB; ++pl) s += tp[(long)pl * S::NCSP] * X[xi(tm.in, m, cls, pl, tm.ic)];
      acc[tm.oc] += tm.sign > 0 ? s : -s;
    }
    out[2 * w] = acc[0];
    out[2 * w + 1] = acc[1];
  }
}

// sum of the latitude-pair blocks' partial sums in a fixed order; n reals
// per block, per_mem of them per member (member stride out_stride in out)
template <typename R>
__global__ void k_sum_blocks(const R* scratch, R* out, int npb, long n, long per_mem,
                             long out_stride) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  R s = 0;
  for (int b = 0; b < npb; ++b) s += scratch[(long)b * n + i];
  const long mem = i / per_mem;
  out[mem * out_stride + (i - mem * per_mem)] = s;
}

template <int T, int MODE>
static int dirx_t(const Tables& tb, const DirxArgs& a, void* scratch, size_t sbytes,
                  hipStream_t s) {
  using C = DirxCfg<T, MODE>;
  using R = RealX<MODE>;
  if (a.nlev <= 0 || a.nmem <= 0) return 0;
  const long per_mem = (long)a.nlev * TS<T>::NRSP;
  const long n = (long)a.nmem * per_mem;
  if (C::NPB > 1 && (!scratch || sbytes < sizeof(R) * n * C::NPB)) return -4;
  const void* tv = (MODE == MODE_F64) ? (const void*)tb.filt64 : (const void*)tb.filt32;
  k_dirx<T, MODE><<<(unsigned)((long)a.nmem * a.nlev * C::NPB), 256, 0, s>>>(a, tv, scratch);
  if (C::NPB > 1)
    k_sum_blocks<R><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(
        (const R*)scratch, (R*)a.out, C::NPB, n, per_mem, a.out_stride);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

size_t dirx_scratch_bytes(const Tables& tb, int mode, int nlev) {
  const long n = (long)nlev * 2 * tb.ncsp;
  const size_t el = (mode == MODE_F64) ? 8 : 4;
  int npb = 1;
  const bool d = (mode == MODE_F64);
  switch (tb.ntru) {
    case 21: npb = d ? DirxCfg<21, MODE_F64>::NPB : DirxCfg<21, MODE_F32>::NPB; break;
    case 42: npb = d ? DirxCfg<42, MODE_F64>::NPB : DirxCfg<42, MODE_F32>::NPB; break;
    case 63: npb = d ? DirxCfg<63, MODE_F64>::NPB : DirxCfg<63, MODE_F32>::NPB; break;
    case 127: npb = d ? DirxCfg<127, MODE_F64>::NPB : DirxCfg<127, MODE_F32>::NPB; break;
  }
  return npb > 1 ? el * n * npb : 0;
}

int launch_dirx(const Tables& tb, int mode, const DirxArgs& a, void* scratch, size_t sbytes,
                hipStream_t s) {
  const bool d = (mode == MODE_F64);
  switch (tb.ntru) {
    case 21: return d ? dirx_t<21, MODE_F64>(tb, a, scratch, sbytes, s)
                      : dirx_t<21, MODE_F32>(tb, a, scratch, sbytes, s);
    case 42: return d ? dirx_t<42, MODE_F64>(tb, a, scratch, sbytes, s)
                      : dirx_t<42, MODE_F32>(tb, a, scratch, sbytes, s);
    case 63: return d ? dirx_t<63, MODE_F64>(tb, a, scratch, sbytes, s)
                      : dirx_t<63, MODE_F32>(tb, a, scratch, sbytes, s);
    case 127: return d ? dirx_t<127, MODE_F64>(tb, a, scratch, sbytes, s)
                       : dirx_t<127, MODE_F32>(tb, a, scratch, sbytes, s);
    default: return -2;
  }
}

}  // namespace exo
