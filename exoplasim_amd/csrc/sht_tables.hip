// sht_tables.hip -- FP64 Gaussian grid, associated-Legendre tables and their
// per-precision packed layouts, all built on the GPU.
//
// k_inigau restates src/gaussmod.f90:5-86 and k_legini restates
// src/legmod.f90:30-116 (same recurrences and evaluation order, FP64).
#include "sht_internal.h"

// no FMA contraction: the recurrences round like the reference's
#pragma clang fp contract(off)

namespace exo {

// ---- Gaussian latitudes: Newton iteration on P_klat (src/gaussmod.f90) ----
__device__ double dev_ql(int k, double p) {
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

__device__ double dev_qld(int k, double p) {
  double z = p * dev_ql(k, p) - sqrt((k + k + 1.0) / (k + k - 1.0)) * dev_ql(k - 1, p);
  return (p * p - 1.0) / (k * z);
}

__global__ void k_inigau(int klat, double* pz0, double* pzw) {
  const int jlat = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (jlat > klat / 2) return;
  const double PI = 3.14159265358979;
  double z0 = PI / (2 * klat + 1);
  double z1 = 1.0 / (double)(klat * klat * 8);
  double z4 = 2.0 / (double)(klat * klat);
  double z2 = z0 * (2 * jlat - 0.5);
  z2 = cos(z2 + z1 / tan(z2));
  for (int it = 0; it < 50; ++it) {
    double z3 = dev_ql(klat, z2) * dev_qld(klat, z2);
    z2 = z2 - z3;
    if (fabs(z3) < 1.0e-16) break;
  }
  double z5 = dev_ql(klat - 1, z2) / sqrt(klat - 0.5);
  pz0[jlat - 1] = z2;
  pzw[jlat - 1] = z4 * (1.0 - z2 * z2) / (z5 * z5);
  pz0[klat - jlat] = -z2;
  pzw[klat - jlat] = pzw[jlat - 1];
}

// ---- Legendre tables, one thread per (north latitude, m) ----
// tab[t][lat][w] for the NPAIR northern latitudes, t in qi..qv
__global__ void k_legini(int ntru, int npair, const double* sid, const double* gwd,
                         double* tab) {
  const int m = blockIdx.y;
  const int lat = blockIdx.x * blockDim.x + threadIdx.x;
  if (lat >= npair || m > ntru) return;
  const int ncsp = (ntru + 1) * (ntru + 2) / 2;
  const long tsz = (long)npair * ncsp;
  const double zgwd = gwd[lat], zsin = sid[lat];
  const double zcsq = 1.0 - zsin * zsin;
  const double zgwdcsq = zgwd / zcsq;
  double P[128], D[128];  // index n - m
  double f1m = sqrt(1.5), f2m = 0.0;
  if (m == 0) {
    P[0] = sqrt(0.5);
    P[1] = f1m * zsin;
    D[0] = 0.0;
  } else {
    for (int mm = 1; mm <= m; ++mm) {
      f2m = -f1m * sqrt(zcsq / (double)(mm + mm));
      f1m = f2m * sqrt(mm + mm + 3.0);
    }
    P[0] = f2m;
    if (m < ntru) {
      P[1] = f1m * zsin;
      D[0] = -(m * f2m * zsin);
    }
  }
  const double amsq = (double)(m * m);
  for (int n = m + 2; n <= ntru; ++n) {
    const int i = n - m;
    double z1 = sqrt(((double)((n - 1) * (n - 1)) - amsq) / (double)(4 * (n - 1) * (n - 1) - 1));
    double z2 = zsin * P[i - 1] - z1 * P[i - 2];
    P[i] = z2 * sqrt((double)(4 * n * n - 1) / ((double)(n * n) - amsq));
    D[i - 1] = (1 - n) * z2 + n * z1 * P[i - 2];
  }
  const int il = ntru - m;
  if (m < ntru) {
    double z3 = sqrt(((double)(ntru * ntru) - amsq) / (double)(4 * ntru * ntru - 1));
    D[il] = -ntru * zsin * P[il] + (ntru + ntru + 1) * P[il - 1] * z3;
  } else {
    D[il] = -ntru * zsin * P[il];
  }
  const int w0 = w0_of(ntru, m);
  for (int n = m; n <= ntru; ++n) {
    const int i = n - m;
    const long o = (long)lat * ncsp + w0 + i;
    double znn1 = 0.0;
    if (n > 0) znn1 = 1.0 / (double)(n * (n + 1));
    tab[TQI * tsz + o] = P[i];
    tab[TQJ * tsz + o] = D[i];
    tab[TQC * tsz + o] = P[i] * zgwd;
    tab[TQU * tsz + o] = P[i] * znn1 * m;
    tab[TQV * tsz + o] = D[i] * znn1;
    tab[TQE * tsz + o] = D[i] * zgwdcsq;
    tab[TQQ * tsz + o] = P[i] * zgwdcsq * n * (n + 1) * 0.5;
    tab[TQM * tsz + o] = P[i] * zgwdcsq * m;
  }
}

// ---- filtered copies used by the VALU paths: tab[t][lat][w] * sk(n) ----
// t in qi, qj, qu, qv uses the SP->GP filter, the others the GP->SP one
__global__ void k_filtered(int ntru, int npair, const double* tab, const double* skgpsp,
                           const double* skspgp, double* out64, float* out32) {
  const int ncsp = (ntru + 1) * (ntru + 2) / 2;
  const long tsz = (long)npair * ncsp;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= NTAB * tsz) return;
  const int t = (int)(i / tsz);
  const int w = (int)(i % ncsp);
  // n of spectral index w
  int m = 0;
  while (m < ntru && w >= w0_of(ntru, m + 1)) ++m;
  const int n = m + (w - w0_of(ntru, m));
  const bool inv = (t == TQI || t == TQJ || t == TQU || t == TQV);
  const double v = tab[i] * (inv ? skspgp[n] : skgpsp[n]);
  out64[i] = v;
  out32[i] = (float)v;
}

// three-way split of an FP64 value into bf16 planes: hi and mid truncated,
// lo rounded to nearest; residual below 2^-24 relative
__device__ void split3_f64(double x, uint16_t& h, uint16_t& m, uint16_t& l) {
  float fx = (float)x;
  uint32_t uh = __float_as_uint(fx) & 0xffff0000u;
  double r1 = x - (double)__uint_as_float(uh);
  uint32_t um = __float_as_uint((float)r1) & 0xffff0000u;
  double r2 = r1 - (double)__uint_as_float(um);
  float f2 = (float)r2;
  uint32_t ul = __float_as_uint(f2);
  ul = (ul + 0x7fffu + ((ul >> 16) & 1u)) & 0xffff0000u;  // round to nearest even
  h = (uint16_t)(uh >> 16);
  m = (uint16_t)(um >> 16);
  l = (uint16_t)(ul >> 16);
}

// ---- table operands of the FP32- and FP64-MFMA inverse: [pb][m][kb][512]
// with MFMA row `row` and k = n - m - 32 kb, stored [k%4][row][k/4] per
// 512-element block so that one lane of v_mfma_f32_16x16x4_f32 (or
// v_mfma_f64_16x16x4_f64, the same operand layout) reads its 8 values
// contiguously; zero where n > NTRU or the class does not match.  Both
// instructions leave D[row][col] in lane 16 q + col, but in item v = row % 4
// of q = row / 4 (FP32) or item v = row / 4 of q = row % 4 (FP64); the rows
// are ordered so that item v of lane q holds pair 2q + v/2, class v % 2 in
// both: FP32 row = 2 pair + class, FP64 row = (row % 4 = pair / 2, row / 4 =
// 2 (pair % 2) + class) ----
__global__ void k_pack_inv(int ntru, int npair, int kbmax, int t, const double* filt64,
                           float* f32, double* f64) {
  const int ncsp = (ntru + 1) * (ntru + 2) / 2;
  const long tsz = (long)npair * ncsp;
  const int npb = npair / 8;
  const long total = (long)npb * (ntru + 1) * kbmax * 512;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int k = (int)(i % 32);
  const int row = (int)((i / 32) % 16);
  const int kb = (int)((i / 512) % kbmax);
  const int m = (int)((i / (512L * kbmax)) % (ntru + 1));
  const int pb = (int)(i / (512L * kbmax * (ntru + 1)));
  const int n = m + kb * 32 + k;
  // the value of (pair pl, class cls) at n
  auto val = [&](int pl, int cls) -> double {
    if (n > ntru) return 0.0;
    const bool even = ((n - m) & 1) == 0;
    const int mycls = (even != table_is_q(t)) ? 0 : 1;
    return mycls == cls ? filt64[t * tsz + (long)(pb * 8 + pl) * ncsp + w0_of(ntru, m) + (n - m)]
                        : 0.0;
  };
  const long blk = i - (i % 512);
  const long o = blk + ((k % 4) * 16 + row) * 8 + k / 4;
  f32[o] = (float)val(row >> 1, row & 1);
  f64[o] = val(2 * (row % 4) + (row / 4) / 2, (row / 4) % 2);
}

// ---- table operand of the split-BF16 inverse (layout in sht_common.h):
// one thread per (pair block, m, row, j) writing the three planes ----
__global__ void k_pack_inv_bf(int ntru, int npair, int t, const double* filt64, uint16_t* bf) {
  const int ncsp = (ntru + 1) * (ntru + 2) / 2;
  const long tsz = (long)npair * ncsp;
  const int per = 16 * pk_c16(ntru + 1);  // (row, j) per pair block
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)(npair / 8) * per) return;
  const int pb = (int)(i / per), e = (int)(i % per);
  int lo = 0, hi = ntru;  // the largest m whose segment starts at or before e
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab_seg(ntru, mid) / 3 <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int m = lo, rw = pk_rw(ntru, m);
  const int r = e - tab_seg(ntru, m) / 3;
  const int row = r / rw, j = r - row * rw;
  const int pl = row >> 1, cls = row & 1;
  const int h = cls ^ (table_is_q(t) ? 1 : 0);
  const int n = m + 2 * j + h;
  const double v =
      (n <= ntru) ? filt64[t * tsz + (long)(pb * 8 + pl) * ncsp + w0_of(ntru, m) + (n - m)] : 0.0;
  uint16_t hh, mm, ll;
  split3_f64(v, hh, mm, ll);
  uint16_t* o = bf + (long)pb * tab_pb(ntru) + tab_seg(ntru, m) + row * rw + j;
  o[0] = hh;
  o[16 * rw] = mm;
  o[32 * rw] = ll;
}

// ---- table operand of the MFMA direct transforms (k_dirx_mf): per (m,
// parity p of n - m, row tile rt, latitude pair) the 16 rows n = m + p +
// 2 (16 rt + r) of one table, [m][p][rt][pair][r], so that the 64 lanes of
// one v_mfma_f32_16x16x4_f32 A operand (16 rows x 4 pairs) read 256
// contiguous bytes; zero where n > NTRU ----
__global__ void k_pack_dir(int ntru, int npair, int rb, int t, const double* filt64, float* out) {
  const int ncsp = (ntru + 1) * (ntru + 2) / 2;
  const long tsz = (long)npair * ncsp;
  const long total = (long)(ntru + 1) * 2 * rb * npair * 16;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int r = (int)(i % 16);
  const int pair = (int)((i / 16) % npair);
  const int rt = (int)((i / (16L * npair)) % rb);
  const int p = (int)((i / (16L * npair * rb)) % 2);
  const int m = (int)(i / (32L * npair * rb));
  const int n = m + p + 2 * (16 * rt + r);
  out[i] = (n <= ntru) ? (float)filt64[t * tsz + (long)pair * ncsp + w0_of(ntru, m) + (n - m)]
                       : 0.0f;
}

// ---- the FP32 tables of the fused direct transform (k_fwd): 4 latitude
// pairs side by side, [t][pair / 4][w][pair % 4], so that a lane summing
// coefficient w reads 4 pairs in one 16-byte load and the lanes of a wave
// (consecutive w) read one contiguous run ----
__global__ void k_pack_fwd4(int npair, int ncsp, const float* filt32, float* fwd4) {
  const long total = (long)NTAB * npair * ncsp;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int u = (int)(i % 4);
  const int w = (int)((i / 4) % ncsp);
  const long tg = i / (4L * ncsp);  // t * (npair / 4) + pair group
  const int g = (int)(tg % (npair / 4));
  const int t = (int)(tg / (npair / 4));
  fwd4[i] = filt32[((long)t * npair + 4 * g + u) * ncsp + w];
}

// twiddles e^{+2 pi i k/N} in FP64 and FP32
__global__ void k_twiddle(int n, double* tw64, float* tw32) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s, c;
  sincospi(2.0 * k / n, &s, &c);
  tw64[2 * k] = c;
  tw64[2 * k + 1] = s;
  tw32[2 * k] = (float)c;
  tw32[2 * k + 1] = (float)s;
}

}  // namespace exo
