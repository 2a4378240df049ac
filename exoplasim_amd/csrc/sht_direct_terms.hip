// sht_direct_terms.hip -- folded direct Legendre transforms with several
// Fourier inputs and tables (src/legmod.f90 folded branches):
//   uv2dv  (:391-422)   (u, v)                       -> (D, zeta)
//   qtend  (:468-496)   (qn, uq, vq)                 -> q
//   mktend (:555-602)   (tn, fu, fv, ke, ut, vt)     -> (D, T, zeta)
// Each output field is a sum of terms  sign * t[pair][w] * X_i[m][cls][pair][c]
// where X_i is input i folded into E = N+S (class sym) / O = N-S (class anti)
// and the class of table t at (m,n) is sym iff (n-m even) != (t is Q-type).
// One launch computes one output field (re and im) for every level; the
// caller launches once per output.  FP64 runs k_dirx (VALU sums, one block
// per level and latitude-pair block); the FP32 modes run k_dirx_mf (MFMA,
// 8 levels per block).
#include <type_traits>

#include "sht_internal.h"

namespace exo {

template <int MODE>
using RealX = typename std::conditional<MODE == MODE_F64, double, float>::type;

// latitude pairs per block so that 3 folded FP64 inputs fit in about
// 100 KB of LDS
template <int T, int MODE>
struct DirxCfg {
  static constexpr int PB = (T == 21) ? 16 : (T == 42) ? 16 : (T == 63) ? 12 : 8;
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

// ---------------------------------------------------------------------------
// the same sums on MFMA (FP32 modes)
// ---------------------------------------------------------------------------
// Per zonal wavenumber m and parity p of n - m, each output is a GEMM over
// the latitude pairs:  C[n][lev, oc] = sum_t sum_pair A_t[n][pair] B_t[pair][lev, oc]
// with A_t the table t at the rows n = m + p + 2j (packed by k_pack_dir) and
// B_t[pair][lev, oc] = sign * X_in[cls][pair][m][lev][ic] for the term of
// table t and output component oc.  One v_mfma_f32_16x16x4_f32 takes 16 rows
// n x 4 pairs of A and 4 pairs x 16 columns (8 levels x re/im) of B.
//
// A block owns one member, 8 levels and a range of MR wavenumbers; it walks
// the latitude pairs KC at a time: the folded Fourier values of those pairs
// (all inputs, its levels and wavenumbers) are staged in LDS, every table
// operand of the step is requested, then the next pairs' values, and the
// MFMAs run while those are in flight; every (m, p, row tile) accumulates in
// registers across all pairs, so the sum over latitudes is complete in the
// block (no partial sums).  Wave w takes m = m0 + w, m0 + w + 8, ...  KC is
// as large as the step's table operands allow in registers: fewer inputs
// (tables), more pairs per step.
template <int T, int NIN>
struct DirxMfCfg {
  static constexpr int MR = 8;  // wavenumbers per block
  static constexpr int NMR = (TS<T>::NM + MR - 1) / MR;
  static constexpr int NT = 512, NW = NT / 64;
  static constexpr int MPW = (MR + NW - 1) / NW;                     // wavenumbers per wave
  static constexpr int LG = 8;                                       // levels per block
  static constexpr int KC = (NIN == 1) ? ((T == 127) ? 8 : 16) : (T == 127) ? ((NIN == 2) ? 8 : 4) : (NIN == 3) ? 8 : (T == 63) ? 16 : 8;  // pairs per step
  static constexpr int KS = LG * 2 * MR + 2;  // floats per (input, class, pair), padded
};

// the term list regrouped by table: slot s multiplies table tab[s] by, for
// output component oc, sign[s][oc] * X_in[s][oc] component ic[s][oc]; every
// term list here has one table per input (mktend, qtend, uv2dv, fc2sp)
constexpr int NSLOT = 3;
struct DirxSlots {
  const float* tab[NSLOT];
  float sign[NSLOT][2];
  int in[NSLOT][2], ic[NSLOT][2];
  int isq[NSLOT];
  int nslot;
};

template <int T, int NIN>
__global__ __launch_bounds__(512) void k_dirx_mf(DirxArgs a, DirxSlots sl) {
  using S = TS<T>;
  using C = DirxMfCfg<T, NIN>;
  constexpr int N = S::NLON, NLAT = S::NLAT, NPAIR = S::NPAIR, RB = S::RB;
  constexpr int MR = C::MR, LG = C::LG, KC = C::KC, KS = C::KS, MPW = C::MPW, NT = C::NT;
  constexpr int NKS = KC / 4;  // MFMA K steps (4 pairs each) per staged step
  constexpr int NJ = 2 * MR;   // reals of one staged row segment
  static_assert(NPAIR % KC == 0, "pairs per step");
  // staged values per step: (input, level, pair, 4 reals) with the reals
  // fastest, one 16-byte load per item
  constexpr int NJ4 = NJ / 4;
  constexpr int NITEM = NIN * LG * KC * NJ4;
  constexpr int MAXIT = (NITEM + NT - 1) / NT;
  __shared__ __align__(16) float X[NIN * 2 * KC * KS];  // [input][class][pair][level][m][re/im]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nlg = (a.nlev + LG - 1) / LG;
  // neighbouring wavenumber ranges stage the two halves of the same 128-byte
  // lines of a Fourier row: renumbered so that they run on one XCD's L2
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int mr = b % C::NMR;
  const int lg = (b / C::NMR) % nlg;
  const int mem = b / (C::NMR * nlg);
  if (mem >= a.nmem) return;
  const int m0 = mr * MR, lev0 = lg * LG;
  const int nj = (S::NM - m0 < MR ? S::NM - m0 : MR) * 2;  // reals below the truncation

  // each input's member as one buffer (2 m0 + NJ <= N: the 16-byte runs
  // stay inside their rows; reals past the truncation are zeroed)
  __amdgpu_buffer_rsrc_t rsi[NIN];
#pragma unroll
  for (int i = 0; i < NIN; ++i)
    rsi[i] = buffer_rsrc((const float*)a.in[i] + mem * a.in_stride, a.nlev * NLAT * N * 4);
  static_assert(2 * ((TS<T>::NM - 1) / MR * MR) + NJ <= N, "staged runs inside a row");
  float rn[MAXIT][4], rs[MAXIT][4];
  auto load = [&](int c) {
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
      const int e = tid + NT * it;
      const int j4 = e % NJ4, k = (e / NJ4) % KC, l = (e / (NJ4 * KC)) % LG,
                i = e / (NJ4 * KC * LG);
#pragma unroll
      for (int q = 0; q < 4; ++q) rn[it][q] = rs[it][q] = 0.f;
      if (e < NITEM && lev0 + l < a.nlev && 4 * j4 < nj) {
        const int pair = c * KC + k;
        const int off = ((lev0 + l) * NLAT * N + 2 * m0 + 4 * j4) * 4;
        const uint4 vn = buffer_load_x4(rsi[i], off + pair * N * 4, 0);
        const uint4 vs = buffer_load_x4(rsi[i], off + (NLAT - 1 - pair) * N * 4, 0);
        const unsigned un[4] = {vn.x, vn.y, vn.z, vn.w}, us[4] = {vs.x, vs.y, vs.z, vs.w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (4 * j4 + q < nj) {
            rn[it][q] = __uint_as_float(un[q]);
            rs[it][q] = __uint_as_float(us[q]);
          }
      }
    }
  };
  // fold into E = N + S (class 0) and O = N - S (class 1)
  auto store = [&]() {
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
      const int e = tid + NT * it;
      if (e < NITEM) {
        const int j4 = e % NJ4, k = (e / NJ4) % KC, l = (e / (NJ4 * KC)) % LG,
                  i = e / (NJ4 * KC * LG);
        const int o = l * NJ + 4 * j4;  // even, and KS is even: 8-byte aligned
        float2* xe = (float2*)&X[((i * 2 + 0) * KC + k) * KS + o];
        float2* xo = (float2*)&X[((i * 2 + 1) * KC + k) * KS + o];
        xe[0] = make_float2(rn[it][0] + rs[it][0], rn[it][1] + rs[it][1]);
        xe[1] = make_float2(rn[it][2] + rs[it][2], rn[it][3] + rs[it][3]);
        xo[0] = make_float2(rn[it][0] - rs[it][0], rn[it][1] - rs[it][1]);
        xo[1] = make_float2(rn[it][2] - rs[it][2], rn[it][3] - rs[it][3]);
      }
    }
  };

  // A operand: row r = lane & 15, pair 4 ks + (lane >> 4) of the step; B
  // operand: column lane & 15 = (level l, component oc), the same pair
  const int r = lane & 15, kq = lane >> 4, l = (lane & 15) >> 1, oc = lane & 1;
  f32x4 acc[MPW][2][RB];
#pragma unroll
  for (int mi = 0; mi < MPW; ++mi)
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int rt = 0; rt < RB; ++rt) acc[mi][p][rt] = f32x4{0.f, 0.f, 0.f, 0.f};
  // this lane's B source per slot: offset of (input, class 0, pair kq,
  // level l, component ic) in X and the sign (0: no term for component oc)
  int boff[NIN];
  float bsg[NIN];
#pragma unroll
  for (int s = 0; s < NIN; ++s) {
    boff[s] = ((oc ? sl.in[s][1] : sl.in[s][0]) * 2 * KC + kq) * KS + l * NJ +
              (oc ? sl.ic[s][1] : sl.ic[s][0]);
    bsg[s] = oc ? sl.sign[s][1] : sl.sign[s][0];
  }
  auto ntile = [](int m, int p) {
    const int d = T - m - p;
    return d < 0 ? 0 : (d / 2 + 16) / 16;
  };

  constexpr int NC = NPAIR / KC;
  load(0);
  for (int c = 0; c < NC; ++c) {
    __syncthreads();  // the previous step's MFMAs have read X
    store();
    __syncthreads();
    // every A operand of this step first (all row tiles: the packed table is
    // zero past the truncation), so that their latencies overlap; then the
    // next step's inputs, whose loads stay in flight through the MFMAs
    float av[MPW][NIN][2][RB][NKS];
#pragma unroll
    for (int mi = 0; mi < MPW; ++mi) {
      const int ml = wv + C::NW * mi, m = m0 + ml;
      if (ml >= MR || m > T) continue;
#pragma unroll
      for (int s = 0; s < NIN; ++s) {
        const float* tp = sl.tab[s] + ((long)(m * 2) * RB * NPAIR + c * KC + kq) * 16 + r;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int rt = 0; rt < RB; ++rt)
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks)
              av[mi][s][p][rt][ks] = tp[(((long)p * RB + rt) * NPAIR + 4 * ks) * 16];
      }
    }
    if (c + 1 < NC) load(c + 1);
#pragma unroll
    for (int mi = 0; mi < MPW; ++mi) {
      const int ml = wv + C::NW * mi, m = m0 + ml;
      if (ml >= MR || m > T) continue;
      const int nt0 = ntile(m, 0), nt1 = ntile(m, 1);
#pragma unroll
      for (int s = 0; s < NIN; ++s) {
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
          float b[2];
#pragma unroll
          for (int cls = 0; cls < 2; ++cls) {
            const float x = X[boff[s] + (cls * KC + 4 * ks) * KS + 2 * ml];
            b[cls] = bsg[s] == 0.f ? 0.f : bsg[s] * x;
          }
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            // class of the folded input: symmetric (E) iff (n - m even) != Q-type table
            const float bv = b[(p == 0) == (sl.isq[s] != 0) ? 1 : 0];
            const int nt = p ? nt1 : nt0;
#pragma unroll
            for (int rt = 0; rt < RB; ++rt)
              if (rt < nt)
                acc[mi][p][rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[mi][s][p][rt][ks], bv,
                                                                      acc[mi][p][rt], 0, 0, 0);
          }
        }
      }
    }
  }
  // lane holds rows 4 kq + v of column (l, oc)
  if (lev0 + l >= a.nlev) return;
  float* out = (float*)a.out + mem * a.out_stride + (long)(lev0 + l) * S::NRSP;
#pragma unroll
  for (int mi = 0; mi < MPW; ++mi) {
    const int ml = wv + C::NW * mi, m = m0 + ml;
    if (ml >= MR || m > T) continue;
    const int w0 = w0_of(T, m);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int rt = 0; rt < RB; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int dn = p + 2 * (16 * rt + 4 * kq + v);
          if (m + dn <= T) out[2 * (w0 + dn) + oc] = acc[mi][p][rt][v];
        }
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

template <int T>
static int dirx_mf_t(const Tables& tb, const DirxArgs& a, hipStream_t s) {
  if (a.nlev <= 0 || a.nmem <= 0) return 0;
  DirxSlots sl{};
  for (int t = 0; t < a.nterm; ++t) {
    const DirxTerm& tm = a.terms[t];
    int k = 0;
    while (k < sl.nslot && sl.tab[k] != tb.dir_f32[tm.table]) ++k;
    if (k == sl.nslot) {
      if (k == NSLOT || !tb.dir_f32[tm.table]) return -2;
      sl.tab[k] = tb.dir_f32[tm.table];
      sl.isq[k] = table_is_q(tm.table) ? 1 : 0;
      ++sl.nslot;
    }
    if (sl.sign[k][tm.oc] != 0.f) return -2;  // one term per (table, component)
    sl.sign[k][tm.oc] = tm.sign > 0 ? 1.f : -1.f;
    sl.in[k][tm.oc] = tm.in;
    sl.ic[k][tm.oc] = tm.ic;
  }
  if (sl.nslot != a.nin) return -2;  // one table per input
  // a member's input levels are addressed as one buffer of 32-bit byte offsets
  if ((long)a.nlev * TS<T>::NLAT * TS<T>::NLON * 4 > 0x7fffffffL) return -2;
  const long nlg = (a.nlev + 7) / 8;
  auto launch = [&](auto kern, int nmr) {
    kern<<<(unsigned)((long)a.nmem * nlg * nmr), 512, 0, s>>>(a, sl);
  };
  switch (a.nin) {
    case 1: launch(k_dirx_mf<T, 1>, DirxMfCfg<T, 1>::NMR); break;
    case 2: launch(k_dirx_mf<T, 2>, DirxMfCfg<T, 2>::NMR); break;
    case 3: launch(k_dirx_mf<T, 3>, DirxMfCfg<T, 3>::NMR); break;
    default: return -2;
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
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
  if (mode != MODE_F64) return 0;  // the MFMA form sums all pairs in one block
  const long n = (long)nlev * 2 * tb.ncsp;
  int npb = 1;
  switch (tb.ntru) {
    case 21: npb = DirxCfg<21, MODE_F64>::NPB; break;
    case 42: npb = DirxCfg<42, MODE_F64>::NPB; break;
    case 63: npb = DirxCfg<63, MODE_F64>::NPB; break;
    case 127: npb = DirxCfg<127, MODE_F64>::NPB; break;
  }
  return npb > 1 ? sizeof(double) * n * npb : 0;
}

int launch_dirx(const Tables& tb, int mode, const DirxArgs& a, void* scratch, size_t sbytes,
                hipStream_t s) {
  // FP64: VALU sums per level; FP32 modes: MFMA over 8 levels per block
  const bool d = (mode == MODE_F64);
  switch (tb.ntru) {
    case 21: return d ? dirx_t<21, MODE_F64>(tb, a, scratch, sbytes, s) : dirx_mf_t<21>(tb, a, s);
    case 42: return d ? dirx_t<42, MODE_F64>(tb, a, scratch, sbytes, s) : dirx_mf_t<42>(tb, a, s);
    case 63: return d ? dirx_t<63, MODE_F64>(tb, a, scratch, sbytes, s) : dirx_mf_t<63>(tb, a, s);
    case 127:
      return d ? dirx_t<127, MODE_F64>(tb, a, scratch, sbytes, s) : dirx_mf_t<127>(tb, a, s);
    default: return -2;
  }
}

}  // namespace exo
