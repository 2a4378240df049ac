// sht_transforms.hip -- the spectral <-> grid transforms on MI355X.
//
//   k_inv  inverse Legendre (src/legmod.f90:230-349: sp2fc, sp2fcdmu,
//          dv2uv), optionally fused with the inverse real FFT
//          (src/fftmod.f90:64-87, fc2gp) so that Fourier coefficients stay
//          in LDS and only spectral coefficients in / grid rows out touch
//          HBM.
//   k_fwd  direct Legendre (src/legmod.f90:178-223: fc2sp, folded branch),
//          optionally fused with the forward real FFT (src/fftmod.f90:29-58,
//          gp2fc).
//   k_fft_rows  the FFTs alone, for the in-place fc2gp/gp2fc entry points.
//
// Legendre arithmetic per precision mode:
//   MODE_F64      FP64 FMA on the vector ALUs (the plain path)
//   MODE_F32      FP32 FMA on the vector ALUs
//   MODE_F32MFMA  v_mfma_f32_16x16x4_f32, FP32 accumulate
//   MODE_BF16X3   v_mfma_f32_16x16x32_bf16 on an exact three-way bf16 split
//                 of both operands, six products (hh, hm, mh, hl, mm, lh),
//                 FP32 accumulate
// FFTs run in FP64 for MODE_F64 and FP32 otherwise.
#include <type_traits>

#include "sht_internal.h"

namespace exo {

template <int MODE>
using RealT = typename std::conditional<MODE == MODE_F64, double, float>::type;

// ---------------------------------------------------------------------------
// tile shapes
// ---------------------------------------------------------------------------
// V selects the tile variant the autotuner times: V = 1 halves the fields
// per block (half the LDS, twice the blocks in flight).
template <int T, int MODE, int V = 0>
struct InvCfg {
  static constexpr int LP = 8;  // latitude pairs per tile (= 16 MFMA rows)
  static constexpr int FB0 =
      (T == 127) ? (MODE == MODE_F64 ? 2 : 4) : (MODE == MODE_F64 ? 4 : 8);
  static constexpr int FB = (FB0 >> V) < 2 ? 2 : (FB0 >> V);
};

template <int T, int MODE, int V = 0>
struct FwdCfg {
  static constexpr int PB = (T == 21) ? 16
                          : (T == 42) ? 32
                          : (T == 63) ? (MODE == MODE_F64 ? 24 : 48)
                                      : (MODE == MODE_F64 ? 16 : 32);
  static constexpr int FBF0 = (T == 21) ? 4 : (T == 42) ? (MODE == MODE_F64 ? 1 : 2) : 1;
  static constexpr int FBF = (FBF0 >> V) < 1 ? 1 : (FBF0 >> V);
  static constexpr int NPB = TS<T>::NPAIR / PB;
};

template <typename R>
struct NS2 {
  C2<R> n, s;  // Fourier coefficient at the northern / southern latitude
};

__device__ __forceinline__ f32x4 mfma_bf16(const BF8& a, const BF8& b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// inverse transform
// ---------------------------------------------------------------------------
template <int T, int MODE, int KIND, bool FUSED, int V>
__global__ __launch_bounds__(256) void k_inv(InvArgs a, const void* tabA0, const void* tabA1,
                                             long plane, const void* tabv, const void* twp) {
  using S = TS<T>;
  using R = RealT<MODE>;
  constexpr int NM = S::NM, N = S::NLON, NLAT = S::NLAT, NPB = S::NPAIR / 8;
  constexpr int LP = InvCfg<T, MODE, V>::LP, FB = InvCfg<T, MODE, V>::FB;
  constexpr int NF = FftBatch<N>::NF;
  constexpr int NTERM = (KIND == INV_UV) ? 2 : 1;

  __shared__ NS2<R> fcl[LP * FB * NM];
  __shared__ C2<R> fbuf[FUSED ? NF * N : 1];
  __shared__ C2<R> tws[FUSED ? N : 1];

  const int tid = threadIdx.x;
  const int ngroups = (KIND == INV_UV) ? (a.nlev + FB / 2 - 1) / (FB / 2) : (a.nfield + FB - 1) / FB;
  const int pb = blockIdx.x % NPB;
  const int g = (blockIdx.x / NPB) % ngroups;
  const int mem = blockIdx.x / (NPB * ngroups);
  if (mem >= a.nmem) return;
  const R* spm = (const R*)a.sp + mem * a.sp_stride;
  R* outm = (R*)a.out + mem * a.out_stride;

  if (FUSED) {
    const C2<R>* tw = (const C2<R>*)twp;
    for (int k = tid; k < N; k += 256) tws[k] = tw[k];
  }

  // value of MFMA/GEMM column c (0..2FB-1) for spectral index w, term t
  auto bval = [&](int term, int c, int w) -> R {
    const int fl = c >> 1, ri = c & 1;
    if (KIND == INV_SCALAR) {
      const int field = g * FB + fl;
      if (fl >= FB || field >= a.nfield) return R(0);
      return spm[(long)(a.src0 + field) * S::NRSP + 2 * w + ri];
    } else {
      const int lvl = g * (FB / 2) + (fl >> 1);
      const int isv = fl & 1;
      if (fl >= FB || lvl >= a.nlev) return R(0);
      const R* zeta = spm + (long)(a.vor0 + lvl) * S::NRSP;
      const R* div = spm + (long)(a.div0 + lvl) * S::NRSP;
      // term 0 (qv): u <- zeta, v <- -D ; term 1 (qu): u <- -iD, v <- -i zeta
      if (term == 0) {
        if (!isv) {
          R z = zeta[2 * w + ri];
          if (w == 1 && ri == 0) z = z - (R)a.plavor;
          return z;
        }
        return -div[2 * w + ri];
      } else {
        const R* x = isv ? zeta : div;
        if (ri == 0) return x[2 * w + 1];
        R xr = x[2 * w];
        if (isv && w == 1) xr = xr - (R)a.plavor;
        return -xr;
      }
    }
  };

  // ---- phase 1: Legendre sums into fcl ----
  if constexpr (MODE == MODE_BF16X3 || MODE == MODE_F32MFMA) {
    // GEMM per m: C[row = 2*pair + class][col] = sum_n A[row][n] B[n][col];
    // wave wv takes m = wv, wv+4, ...; each (m, 32-wide n block) is one
    // work item, and the operands of the next item are fetched before the
    // current one is multiplied.
    const int lane = tid & 63, wv = tid >> 6;
    const int q = lane >> 4, c = lane & 15;
    // column c of B: SCALAR field g*FB + c/2; UV level g*FB/2 + c/4 of
    // vorticity ((c/2)&1 == 0) or divergence; component c&1
    const R* bp = nullptr;
    bool isvor = false;
    if (KIND == INV_SCALAR) {
      const int field = g * FB + (c >> 1);
      if ((c >> 1) < FB && field < a.nfield) bp = spm + (long)(a.src0 + field) * S::NRSP + (c & 1);
    } else {
      const int lvl = g * (FB / 2) + (c >> 2);
      const int s = (c >> 1) & 1;
      if ((c >> 2) < FB / 2 && lvl < a.nlev) {
        bp = spm + (long)((s ? a.div0 : a.vor0) + lvl) * S::NRSP + (c & 1);
        isvor = (s == 0) && ((c & 1) == 0);
      }
    }
    const R pv = (R)a.plavor;
    // n offset of B element j (bf16: 8 consecutive n per lane; f32: n = 4j + q)
    auto noff = [&](int j) { return MODE == MODE_BF16X3 ? q * 8 + j : j * 4 + q; };
    auto fetch = [&](int m, int kb, float* br, uint4* a0, uint4* a1) {
      const int w0 = w0_of(T, m);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int n = m + kb * 32 + noff(j);
        R x = R(0);
        if (bp && n <= T) {
          x = bp[2 * (w0 + n - m)];
          if (isvor && m == 0 && n == 1) x = x - pv;
        }
        br[j] = (float)x;
      }
      const long blk = ((long)(pb * NM + m) * S::KB + kb) * 512;
      if constexpr (MODE == MODE_BF16X3) {
        const uint16_t* A0 = (const uint16_t*)tabA0 + blk + c * 32 + q * 8;
#pragma unroll
        for (int p = 0; p < 3; ++p) a0[p] = *(const uint4*)(A0 + p * plane);
        if (NTERM == 2) {
          const uint16_t* A1 = (const uint16_t*)tabA1 + blk + c * 32 + q * 8;
#pragma unroll
          for (int p = 0; p < 3; ++p) a1[p] = *(const uint4*)(A1 + p * plane);
        }
      } else {
        // f32 operand stored as [q][row][8] so one lane reads 8 floats
        const float* A0 = (const float*)tabA0 + blk + (q * 16 + c) * 8;
        a0[0] = *(const uint4*)A0;
        a0[1] = *(const uint4*)(A0 + 4);
        if (NTERM == 2) {
          const float* A1 = (const float*)tabA1 + blk + (q * 16 + c) * 8;
          a1[0] = *(const uint4*)A1;
          a1[1] = *(const uint4*)(A1 + 4);
        }
      }
    };
    auto mul = [&](const float* br, const uint4* a, f32x4 acc) -> f32x4 {
      if constexpr (MODE == MODE_BF16X3) {
        BF8 bh, bm, bl, ah, am, al;
#pragma unroll
        for (int j = 0; j < 8; ++j) split3(br[j], bh.s[j], bm.s[j], bl.s[j]);
        ah.q = a[0];
        am.q = a[1];
        al.q = a[2];
        acc = mfma_bf16(ah, bl, acc);
        acc = mfma_bf16(am, bm, acc);
        acc = mfma_bf16(al, bh, acc);
        acc = mfma_bf16(ah, bm, acc);
        acc = mfma_bf16(am, bh, acc);
        acc = mfma_bf16(ah, bh, acc);
      } else {
        const float* af = (const float*)a;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[j], br[j], acc, 0, 0, 0);
      }
      return acc;
    };
    float br[8], br2[8];
    uint4 a0[3], a1[3], a02[3], a12[3];
    int m = wv, kb = 0;
    if (m <= T) fetch(m, 0, br, a0, a1);
    f32x4 acc1 = {0.f, 0.f, 0.f, 0.f}, acc2 = {0.f, 0.f, 0.f, 0.f};
    while (m <= T) {
      const int kbn = (NM - m + 31) / 32;
      int m2 = m, kb2 = kb + 1;
      if (kb2 >= kbn) {
        m2 = m + 4;
        kb2 = 0;
      }
      if (m2 <= T) fetch(m2, kb2, br2, a02, a12);
      acc1 = mul(br, a0, acc1);
      if (NTERM == 2) acc2 = mul(br, a1, acc2);
      if (kb2 == 0) {
        // all n of this m are summed: combine the terms, fold N/S, store
        f32x4 out = acc1;
        if (KIND == INV_UV) {
          // u = qv.zeta + qu.(-i D), v = qu.(-i zeta) - qv.D: the qu product of
          // the partner column (zeta.re <-> D.im, zeta.im <-> D.re)
          const int plane_ = (lane & ~3) | (3 - (lane & 3));
          f32x4 p2;
#pragma unroll
          for (int i = 0; i < 4; ++i) p2[i] = __shfl(acc2[i], plane_, 64);
          const int r4 = c & 3;
#pragma unroll
          for (int i = 0; i < 4; ++i)
            out[i] = (r4 == 0) ? acc1[i] + p2[i]
                   : (r4 == 1) ? acc1[i] - p2[i]
                   : (r4 == 2) ? p2[i] - acc1[i]
                               : -p2[i] - acc1[i];
        }
        const int fl = c >> 1, ri = c & 1;
        if (fl < FB) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int pl = 2 * q + h;
            const float sy = out[2 * h], an = out[2 * h + 1];
            NS2<R>& e = fcl[(pl * FB + fl) * NM + m];
            if (ri == 0) {
              e.n.x = sy + an;
              e.s.x = sy - an;
            } else {
              e.n.y = sy + an;
              e.s.y = sy - an;
            }
          }
        }
        acc1 = f32x4{0.f, 0.f, 0.f, 0.f};
        acc2 = f32x4{0.f, 0.f, 0.f, 0.f};
      }
      m = m2;
      kb = kb2;
#pragma unroll
      for (int j = 0; j < 8; ++j) br[j] = br2[j];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        a0[p] = a02[p];
        a1[p] = a12[p];
      }
    }
  } else {
    // vector-ALU Legendre: one (pair, column, m) sum per thread
    const R* tv = (const R*)tabv;
    const long tsz = (long)S::NPAIR * S::NCSP;
    constexpr int NPC = LP * 2 * FB;
    const int t0 = (KIND == INV_UV) ? TQV : a.table;
    const int t1 = TQU;
    for (int idx = tid; idx < NPC * NM; idx += 256) {
      const int pc = idx % NPC, m = idx / NPC;
      const int pl = pc % LP, c = pc / LP;
      const int pair = pb * LP + pl;
      const int w0 = w0_of(T, m);
      R sy = 0, an = 0;
      for (int n = m; n <= T; ++n) {
        const int w = w0 + n - m;
        const bool even = ((n - m) & 1) == 0;
#pragma unroll
        for (int term = 0; term < NTERM; ++term) {
          const int t = term == 0 ? t0 : t1;
          const R p = tv[t * tsz + (long)pair * S::NCSP + w];
          const R x = bval(term, c, w);
          if (even != table_is_q(t))
            sy += p * x;
          else
            an += p * x;
        }
      }
      const int fl = c >> 1, ri = c & 1;
      NS2<R>& e = fcl[(pl * FB + fl) * NM + m];
      if (ri == 0) {
        e.n.x = sy + an;
        e.s.x = sy - an;
      } else {
        e.n.y = sy + an;
        e.s.y = sy - an;
      }
    }
  }
  __syncthreads();

  // destination field of FFT-field fl (-1 if not present)
  auto dst_of = [&](int fl) -> int {
    if (KIND == INV_SCALAR) {
      const int field = g * FB + fl;
      return field < a.nfield ? a.dst0 + field : -1;
    }
    const int lvl = g * (FB / 2) + (fl >> 1);
    if (lvl >= a.nlev) return -1;
    return ((fl & 1) ? a.v0 : a.u0) + lvl;
  };

  if constexpr (!FUSED) {
    // Fourier output fc(2, NLON/2, NLAT), zero above the truncation
    for (int idx = tid; idx < LP * FB * (N / 2); idx += 256) {
      const int m = idx % (N / 2), pf = idx / (N / 2);
      const int fl = pf % FB, pl = pf / FB;
      const int d = dst_of(fl);
      if (d < 0) continue;
      const int lat = pb * LP + pl;
      C2<R> vn = {0, 0}, vs = {0, 0};
      if (m <= T) {
        vn = fcl[(pl * FB + fl) * NM + m].n;
        vs = fcl[(pl * FB + fl) * NM + m].s;
      }
      R* o = outm + (long)d * NLAT * N;
      o[(long)lat * N + 2 * m] = vn.x;
      o[(long)lat * N + 2 * m + 1] = vn.y;
      o[(long)(NLAT - 1 - lat) * N + 2 * m] = vs.x;
      o[(long)(NLAT - 1 - lat) * N + 2 * m + 1] = vs.y;
    }
  } else {
    // ---- phase 2: inverse FFT of z = fc_N + i fc_S, Re -> north row,
    // Im -> south row (imag of m=0 ignored, as fc2gp does) ----
    for (int f0 = 0; f0 < LP * FB; f0 += NF) {
      const int nf = min(NF, LP * FB - f0);
      auto load = [&](int f, int k) -> C2<R> {
        const int gi = f0 + f, pl = gi / FB, fl = gi % FB;
        const NS2<R>* e = fcl + (pl * FB + fl) * NM;
        if (k == 0) return C2<R>{e[0].n.x, e[0].s.x};
        if (k <= T) {
          const NS2<R> v = e[k];
          return C2<R>{v.n.x - v.s.y, v.n.y + v.s.x};
        }
        if (k >= N - T) {
          const NS2<R> v = e[N - k];
          return C2<R>{v.n.x + v.s.y, v.s.x - v.n.y};
        }
        return C2<R>{R(0), R(0)};
      };
      auto store = [&](int f, int k, C2<R> v) {
        const int gi = f0 + f, pl = gi / FB, fl = gi % FB;
        const int d = dst_of(fl);
        if (d < 0) return;
        const int lat = pb * LP + pl;
        R* o = outm + (long)d * NLAT * N;
        o[(long)lat * N + k] = v.x;
        o[(long)(NLAT - 1 - lat) * N + k] = v.y;
      };
      fft_run<N, +1, R>(fbuf, tws, nf, load, store);
    }
  }
}

// ---------------------------------------------------------------------------
// direct transform
// ---------------------------------------------------------------------------
// One block transforms FBF field-levels of one member over PB latitude pairs
// (all pairs when NPB == 1, else partial sums reduced by k_reduce_partials).
// The direct Legendre sum runs on the vector ALUs in every mode: a block can
// hold the Fourier coefficients of only FBF (1-4) fields in LDS, which would
// fill 2-8 of the 16 columns of an MFMA tile while the operand split would
// still cost the whole wave (see the design notes).
template <int T, int MODE, bool FUSED, int V>
__global__ __launch_bounds__(256) void k_fwd(FwdArgs a, const void* tabv, const void* twp,
                                             void* scratch) {
  using S = TS<T>;
  using R = RealT<MODE>;
  using C = FwdCfg<T, MODE, V>;
  constexpr int NM = S::NM, N = S::NLON, NLAT = S::NLAT;
  constexpr int PB = C::PB, FBF = C::FBF, NPB = C::NPB;
  constexpr int NCOL = 2 * FBF;
  constexpr int NF = FftBatch<N>::NF;

  // F[m][cls][pair][col]: cls 0 = fc_N + fc_S (E), cls 1 = fc_N - fc_S (O)
  __shared__ R F[NM * 2 * PB * NCOL];
  __shared__ C2<R> fbuf[FUSED ? NF * N : 1];
  __shared__ C2<R> tws[FUSED ? N : 1];
  __shared__ uint8_t mw[S::NCSP];

  const int tid = threadIdx.x;
  const int ngroups = (a.nfield + FBF - 1) / FBF;
  const int pb = blockIdx.x % NPB;
  const int g = (blockIdx.x / NPB) % ngroups;
  const int mem = blockIdx.x / (NPB * ngroups);
  if (mem >= a.nmem) return;
  const R* inm = (const R*)a.in + mem * a.in_stride;

  for (int m = tid; m <= T; m += 256)
    for (int n = m; n <= T; ++n) mw[w0_of(T, m) + n - m] = (uint8_t)m;
  if (FUSED) {
    const C2<R>* tw = (const C2<R>*)twp;
    for (int k = tid; k < N; k += 256) tws[k] = C2<R>{tw[k].x, -tw[k].y};
  }
  __syncthreads();

  auto fidx = [](int m, int cls, int pl, int col) { return ((m * 2 + cls) * PB + pl) * NCOL + col; };

  if constexpr (FUSED) {
    for (int f0 = 0; f0 < PB * FBF; f0 += NF) {
      const int nf = min(NF, PB * FBF - f0);
      auto load = [&](int f, int k) -> C2<R> {
        const int gi = f0 + f, pl = gi / FBF, fl = gi % FBF;
        const int field = g * FBF + fl;
        if (field >= a.nfield) return C2<R>{R(0), R(0)};
        const int lat = pb * PB + pl;
        const R* src = inm + (long)(a.src0 + field) * NLAT * N;
        return C2<R>{src[(long)lat * N + k], src[(long)(NLAT - 1 - lat) * N + k]};
      };
      auto store = [&](int f, int k, C2<R> v) { fbuf[f * N + k] = v; };
      fft_run<N, -1, R>(fbuf, tws, nf, load, store);
      // separate the two real rows, scale by 1/N, fold into E/O
      const R sc = R(0.5) / R(N);
      for (int idx = tid; idx < nf * NM; idx += 256) {
        const int f = idx % nf, k = idx / nf;
        const C2<R> zk = fbuf[f * N + k];
        const C2<R> zn = fbuf[f * N + ((N - k) % N)];
        const R cnx = (zk.x + zn.x) * sc, cny = (zk.y - zn.y) * sc;
        const R csx = (zk.y + zn.y) * sc, csy = (zn.x - zk.x) * sc;
        const int gi = f0 + f, pl = gi / FBF, fl = gi % FBF;
        F[fidx(k, 0, pl, 2 * fl)] = cnx + csx;
        F[fidx(k, 0, pl, 2 * fl + 1)] = cny + csy;
        F[fidx(k, 1, pl, 2 * fl)] = cnx - csx;
        F[fidx(k, 1, pl, 2 * fl + 1)] = cny - csy;
      }
      __syncthreads();
    }
  } else {
    for (int idx = tid; idx < PB * FBF * NM; idx += 256) {
      const int pf = idx % (PB * FBF), k = idx / (PB * FBF);
      const int fl = pf % FBF, pl = pf / FBF;
      const int field = g * FBF + fl;
      if (field >= a.nfield) continue;
      const int lat = pb * PB + pl;
      const R* src = inm + (long)(a.src0 + field) * NLAT * N;
      const R nx = src[(long)lat * N + 2 * k], ny = src[(long)lat * N + 2 * k + 1];
      const R sx = src[(long)(NLAT - 1 - lat) * N + 2 * k];
      const R sy = src[(long)(NLAT - 1 - lat) * N + 2 * k + 1];
      F[fidx(k, 0, pl, 2 * fl)] = nx + sx;
      F[fidx(k, 0, pl, 2 * fl + 1)] = ny + sy;
      F[fidx(k, 1, pl, 2 * fl)] = nx - sx;
      F[fidx(k, 1, pl, 2 * fl + 1)] = ny - sy;
    }
    __syncthreads();
  }

  R* outp;
  if (NPB == 1)
    outp = (R*)a.sp + mem * a.sp_stride + (long)a.dst0 * S::NRSP;
  else
    outp = (R*)scratch + ((long)pb * a.nmem + mem) * a.nfield * S::NRSP;
  const bool qtab = table_is_q(a.table);
  const R* tv = (const R*)tabv + (long)a.table * S::NPAIR * S::NCSP + (long)pb * PB * S::NCSP;
  for (int w = tid; w < S::NCSP; w += 256) {
    const int m = mw[w];
    const int n = m + (w - w0_of(T, m));
    const int cls = ((((n - m) & 1) == 0) != qtab) ? 0 : 1;
    const R* fr = &F[fidx(m, cls, 0, 0)];
    R acc[NCOL];
#pragma unroll
    for (int c = 0; c < NCOL; ++c) acc[c] = R(0);
#pragma unroll 4
    for (int pl = 0; pl < PB; ++pl) {
      const R p = tv[(long)pl * S::NCSP + w];
#pragma unroll
      for (int c = 0; c < NCOL; ++c) acc[c] += p * fr[pl * NCOL + c];
    }
#pragma unroll
    for (int c = 0; c < NCOL; ++c) {
      const int field = g * FBF + (c >> 1);
      if (field < a.nfield) outp[(long)field * S::NRSP + 2 * w + (c & 1)] = acc[c];
    }
  }
}

// sum of the per-pair-block partial direct transforms in a fixed order
template <typename R>
__global__ void k_reduce_partials(const R* scratch, R* sp, int npb, int nmem, int nfield,
                                  int nrsp, long sp_stride, int dst0) {
  const long total = (long)nmem * nfield * nrsp;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int x = (int)(i % nrsp);
  const int f = (int)((i / nrsp) % nfield);
  const int mem = (int)(i / ((long)nrsp * nfield));
  R s = 0;
  for (int b = 0; b < npb; ++b) s += scratch[(long)b * total + i];
  sp[mem * sp_stride + (long)(dst0 + f) * nrsp + x] = s;
}

// ---------------------------------------------------------------------------
// FFT of independent rows (in place), two real rows per complex FFT
// DIR +1: Fourier -> grid (fc2gp), DIR -1: grid -> Fourier (gp2fc)
// ---------------------------------------------------------------------------
template <int N, typename R, int DIR>
__global__ __launch_bounds__(256) void k_fft_rows(R* a, long lot, const C2<R>* tw) {
  constexpr int NF = FftBatch<N>::NF;
  __shared__ C2<R> fbuf[NF * N];
  __shared__ C2<R> tws[N];
  for (int k = threadIdx.x; k < N; k += 256)
    tws[k] = DIR > 0 ? tw[k] : C2<R>{tw[k].x, -tw[k].y};
  __syncthreads();
  const long r0 = (long)blockIdx.x * 2 * NF;
  const long nrow = lot - r0 < 2 * NF ? lot - r0 : 2 * NF;
  const int nf = (int)((nrow + 1) / 2);
  auto row = [&](int f, int h) -> R* {
    const long r = r0 + 2 * f + h;
    return r < lot ? a + r * N : nullptr;
  };
  if (DIR > 0) {
    auto load = [&](int f, int k) -> C2<R> {
      const R* x0 = row(f, 0);
      const R* x1 = row(f, 1);
      auto cf = [&](const R* x, int kk) -> C2<R> {
        if (!x) return C2<R>{R(0), R(0)};
        return C2<R>{x[2 * kk], kk == 0 ? R(0) : x[2 * kk + 1]};
      };
      if (k == N / 2) return C2<R>{R(0), R(0)};
      if (k < N / 2) {
        const C2<R> c0 = cf(x0, k), c1 = cf(x1, k);
        return C2<R>{c0.x - c1.y, c0.y + c1.x};
      }
      const C2<R> c0 = cf(x0, N - k), c1 = cf(x1, N - k);
      return C2<R>{c0.x + c1.y, c1.x - c0.y};
    };
    auto store = [&](int f, int k, C2<R> v) {
      R* x0 = row(f, 0);
      R* x1 = row(f, 1);
      if (x0) x0[k] = v.x;
      if (x1) x1[k] = v.y;
    };
    fft_run<N, +1, R>(fbuf, tws, nf, load, store);
  } else {
    auto load = [&](int f, int k) -> C2<R> {
      const R* x0 = row(f, 0);
      const R* x1 = row(f, 1);
      return C2<R>{x0 ? x0[k] : R(0), x1 ? x1[k] : R(0)};
    };
    auto store = [&](int f, int k, C2<R> v) { fbuf[f * N + k] = v; };
    fft_run<N, -1, R>(fbuf, tws, nf, load, store);
    const R sc = R(0.5) / R(N);
    for (int idx = threadIdx.x; idx < nf * (N / 2); idx += 256) {
      const int f = idx / (N / 2), k = idx % (N / 2);
      const C2<R> zk = fbuf[f * N + k];
      const C2<R> zn = fbuf[f * N + ((N - k) % N)];
      R* x0 = row(f, 0);
      R* x1 = row(f, 1);
      if (x0) {
        x0[2 * k] = (zk.x + zn.x) * sc;
        x0[2 * k + 1] = (zk.y - zn.y) * sc;
      }
      if (x1) {
        x1[2 * k] = (zk.y + zn.y) * sc;
        x1[2 * k + 1] = (zn.x - zk.x) * sc;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int T, int MODE, int V>
static int inv_tv(const Tables& tb, bool fused, const InvArgs& a, hipStream_t s) {
  using C = InvCfg<T, MODE, V>;
  constexpr int NPB = TS<T>::NPAIR / 8;
  const int ngroups = (a.kind == INV_UV) ? (a.nlev + C::FB / 2 - 1) / (C::FB / 2)
                                         : (a.nfield + C::FB - 1) / C::FB;
  const long nblk = (long)a.nmem * ngroups * NPB;
  if (nblk <= 0) return 0;
  const void *A0 = nullptr, *A1 = nullptr;
  const int t0 = (a.kind == INV_UV) ? TQV : a.table;
  if (MODE == MODE_BF16X3) {
    A0 = tb.inv_bf[t0];
    A1 = tb.inv_bf[TQU];
  } else if (MODE == MODE_F32MFMA) {
    A0 = tb.inv_f32[t0];
    A1 = tb.inv_f32[TQU];
  }
  const void* tv = (MODE == MODE_F64) ? (const void*)tb.filt64 : (const void*)tb.filt32;
  const void* tw = (MODE == MODE_F64) ? (const void*)tb.tw64 : (const void*)tb.tw32;
  dim3 grid((unsigned)nblk), block(256);
  if (a.kind == INV_UV) {
    if (fused)
      k_inv<T, MODE, INV_UV, true, V><<<grid, block, 0, s>>>(a, A0, A1, tb.inv_elems, tv, tw);
    else
      k_inv<T, MODE, INV_UV, false, V><<<grid, block, 0, s>>>(a, A0, A1, tb.inv_elems, tv, tw);
  } else {
    if (fused)
      k_inv<T, MODE, INV_SCALAR, true, V><<<grid, block, 0, s>>>(a, A0, A1, tb.inv_elems, tv, tw);
    else
      k_inv<T, MODE, INV_SCALAR, false, V><<<grid, block, 0, s>>>(a, A0, A1, tb.inv_elems, tv, tw);
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// the unfused (drop-in) transforms always use the default tiles
template <int T, int MODE>
static int inv_t(const Tables& tb, bool fused, const InvArgs& a, hipStream_t s) {
  if (fused && a.var == 1) return inv_tv<T, MODE, 1>(tb, fused, a, s);
  return inv_tv<T, MODE, 0>(tb, fused, a, s);
}

template <int T, int MODE>
static size_t fwd_scratch_t(int nmem, int nfield) {
  using C = FwdCfg<T, MODE>;
  if (C::NPB == 1) return 0;
  return (size_t)C::NPB * nmem * nfield * TS<T>::NRSP * sizeof(RealT<MODE>);
}

template <int T, int MODE, int V>
static int fwd_tv(const Tables& tb, bool fused, const FwdArgs& a, void* scratch, size_t sbytes,
                  hipStream_t s) {
  using C = FwdCfg<T, MODE, V>;
  using R = RealT<MODE>;
  const int ngroups = (a.nfield + C::FBF - 1) / C::FBF;
  const long nblk = (long)a.nmem * ngroups * C::NPB;
  if (nblk <= 0) return 0;
  if (C::NPB > 1 && (scratch == nullptr || sbytes < fwd_scratch_t<T, MODE>(a.nmem, a.nfield)))
    return -4;
  const void* tv = (MODE == MODE_F64) ? (const void*)tb.filt64 : (const void*)tb.filt32;
  const void* tw = (MODE == MODE_F64) ? (const void*)tb.tw64 : (const void*)tb.tw32;
  dim3 grid((unsigned)nblk), block(256);
  if (fused)
    k_fwd<T, MODE, true, V><<<grid, block, 0, s>>>(a, tv, tw, scratch);
  else
    k_fwd<T, MODE, false, V><<<grid, block, 0, s>>>(a, tv, tw, scratch);
  if (C::NPB > 1) {
    const long total = (long)a.nmem * a.nfield * TS<T>::NRSP;
    k_reduce_partials<R><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(
        (const R*)scratch, (R*)a.sp, C::NPB, a.nmem, a.nfield, TS<T>::NRSP, a.sp_stride, a.dst0);
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <int T, int MODE>
static int fwd_t(const Tables& tb, bool fused, const FwdArgs& a, void* scratch, size_t sbytes,
                 hipStream_t s) {
  if (fused && a.var == 1) return fwd_tv<T, MODE, 1>(tb, fused, a, scratch, sbytes, s);
  return fwd_tv<T, MODE, 0>(tb, fused, a, scratch, sbytes, s);
}

#define EXO_DISPATCH_T(fn, ...)                                   \
  switch (tb.ntru) {                                              \
    case 21: return fn<21, MODE_X>(__VA_ARGS__);                  \
    case 42: return fn<42, MODE_X>(__VA_ARGS__);                  \
    case 63: return fn<63, MODE_X>(__VA_ARGS__);                  \
    case 127: return fn<127, MODE_X>(__VA_ARGS__);                \
    default: return -2;                                           \
  }

template <int MODE_X>
static int inv_mode(const Tables& tb, bool fused, const InvArgs& a, hipStream_t s) {
  EXO_DISPATCH_T(inv_t, tb, fused, a, s)
}
template <int MODE_X>
static int fwd_mode(const Tables& tb, bool fused, const FwdArgs& a, void* scr, size_t sb,
                    hipStream_t s) {
  EXO_DISPATCH_T(fwd_t, tb, fused, a, scr, sb, s)
}
template <int MODE_X>
static size_t scr_mode(const Tables& tb, int nmem, int nfield) {
  switch (tb.ntru) {
    case 21: return fwd_scratch_t<21, MODE_X>(nmem, nfield);
    case 42: return fwd_scratch_t<42, MODE_X>(nmem, nfield);
    case 63: return fwd_scratch_t<63, MODE_X>(nmem, nfield);
    case 127: return fwd_scratch_t<127, MODE_X>(nmem, nfield);
    default: return 0;
  }
}

int launch_inverse(const Tables& tb, int mode, bool fused, const InvArgs& a, hipStream_t s) {
  switch (mode) {
    case MODE_F64: return inv_mode<MODE_F64>(tb, fused, a, s);
    case MODE_F32: return inv_mode<MODE_F32>(tb, fused, a, s);
    case MODE_F32MFMA: return inv_mode<MODE_F32MFMA>(tb, fused, a, s);
    case MODE_BF16X3: return inv_mode<MODE_BF16X3>(tb, fused, a, s);
    default: return -1;
  }
}

int launch_forward(const Tables& tb, int mode, bool fused, const FwdArgs& a, void* scratch,
                   size_t sbytes, hipStream_t s) {
  switch (mode) {
    case MODE_F64: return fwd_mode<MODE_F64>(tb, fused, a, scratch, sbytes, s);
    case MODE_F32: return fwd_mode<MODE_F32>(tb, fused, a, scratch, sbytes, s);
    case MODE_F32MFMA: return fwd_mode<MODE_F32MFMA>(tb, fused, a, scratch, sbytes, s);
    case MODE_BF16X3: return fwd_mode<MODE_BF16X3>(tb, fused, a, scratch, sbytes, s);
    default: return -1;
  }
}

size_t forward_scratch_bytes(const Tables& tb, int mode, int nmem, int nfield) {
  switch (mode) {
    case MODE_F64: return scr_mode<MODE_F64>(tb, nmem, nfield);
    case MODE_F32: return scr_mode<MODE_F32>(tb, nmem, nfield);
    case MODE_F32MFMA: return scr_mode<MODE_F32MFMA>(tb, nmem, nfield);
    case MODE_BF16X3: return scr_mode<MODE_BF16X3>(tb, nmem, nfield);
    default: return 0;
  }
}

template <int N, typename R>
static int fft_t(const void* tw, int dir, R* a, long lot, hipStream_t s) {
  constexpr int NF = FftBatch<N>::NF;
  const long nblk = (lot + 2 * NF - 1) / (2 * NF);
  if (nblk <= 0) return 0;
  if (dir > 0)
    k_fft_rows<N, R, +1><<<(unsigned)nblk, 256, 0, s>>>(a, lot, (const C2<R>*)tw);
  else
    k_fft_rows<N, R, -1><<<(unsigned)nblk, 256, 0, s>>>(a, lot, (const C2<R>*)tw);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_fft(const Tables& tb, int mode, int dir, void* a, int n, long lot, hipStream_t s) {
  if (n != tb.nlon) return -2;
  const bool d = (mode == MODE_F64);
  switch (n) {
    case 64: return d ? fft_t<64, double>(tb.tw64, dir, (double*)a, lot, s)
                      : fft_t<64, float>(tb.tw32, dir, (float*)a, lot, s);
    case 128: return d ? fft_t<128, double>(tb.tw64, dir, (double*)a, lot, s)
                       : fft_t<128, float>(tb.tw32, dir, (float*)a, lot, s);
    case 192: return d ? fft_t<192, double>(tb.tw64, dir, (double*)a, lot, s)
                       : fft_t<192, float>(tb.tw32, dir, (float*)a, lot, s);
    case 384: return d ? fft_t<384, double>(tb.tw64, dir, (double*)a, lot, s)
                       : fft_t<384, float>(tb.tw32, dir, (float*)a, lot, s);
    default: return -2;
  }
}

}  // namespace exo
