// sht_transforms_impl.h -- the per-truncation transform kernels and their
// launchers for one precision mode.  Each sht_transforms_m<mode>.hip
// includes this file and instantiates one mode, so the four modes compile in
// parallel; sht_transforms.hip dispatches on the mode.
#pragma once
#include "sht_bodies.h"
#include "sht_launch.h"

namespace exo {

template <int T, int MODE, int KIND, bool FUSED, int V>
__global__ __launch_bounds__((InvCfg<T, MODE, V>::NT)) void k_inv(InvArgs a, const void* tabA0, const void* tabA1,
                                             long plane, const void* tabv, const void* twp) {
  __shared__ __align__(16) char lds[InvLds<T, MODE, FUSED, V>::BYTES];
  inv_body<T, MODE, KIND, FUSED, V>(a, xcd_block(blockIdx.x, gridDim.x), tabA0, tabA1, plane,
                                    tabv, twp, lds);
}

// The three inverse transforms of a round trip in one launch: per member,
// the sp2fc blocks (aq, table qi), the sp2fcdmu blocks (aj, table qj) and the
// dv2uv blocks (au), each block the same work item as in its own kernel.
// One launch instead of three removes two ramp-up/drain tails, and the
// blocks of one member (which read the same spectral coefficients) run
// close together.
template <int T, int MODE, int V>
__global__ __launch_bounds__((InvCfg<T, MODE, V>::NT)) void k_inv_rt(InvArgs aq, InvArgs aj, InvArgs au, const void* tq,
                                                const void* tj, const void* tv0, const void* tu,
                                                long plane, const void* tabv, const void* twp,
                                                int gq, int gj, int gu) {
  __shared__ __align__(16) char lds[InvLds<T, MODE, true, V>::BYTES];
  constexpr int NPB = TS<T>::NPAIR / InvCfg<T, MODE, V>::LP;
  const int per = NPB * (gq + gj + gu);
  int mem, r;
  if (T == 127 && aq.nmem % 8 == 0) {
    // T127: workgroups go round-robin to the 8 XCDs, and XCD x =
    // blockIdx.x % 8 runs members x M .. x M + M - 1 (M = nmem / 8), first
    // all their dv2uv items, then all their scalar items: the round trip of
    // 8 members at L40 1% faster than the member-major order below (which
    // stays faster at T42, by about as much)
    const int x = blockIdx.x % 8, i = blockIdx.x / 8, M = aq.nmem / 8;
    const int nu = M * NPB * gu, ns = NPB * (gq + gj);
    if (i < nu) {
      mem = x * M + i / (NPB * gu);
      r = ns + i % (NPB * gu);
    } else {
      mem = x * M + (i - nu) / ns;
      r = (i - nu) % ns;
    }
  } else {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    mem = b / per;
    r = b - mem * per;
  }
  if (r < NPB * gq)
    inv_body<T, MODE, INV_SCALAR, true, V>(aq, mem * NPB * gq + r, tq, tu, plane, tabv, twp, lds);
  else if (r < NPB * (gq + gj))
    inv_body<T, MODE, INV_SCALAR, true, V>(aj, mem * NPB * gj + (r - NPB * gq), tj, tu, plane,
                                           tabv, twp, lds);
  else
    inv_body<T, MODE, INV_UV, true, V>(au, mem * NPB * gu + (r - NPB * (gq + gj)), tv0, tu, plane,
                                       tabv, twp, lds);
}

template <int T, int MODE, bool FUSED, int V>
__global__ __launch_bounds__((FwdCfg<T, MODE, V>::NT)) void k_fwd(FwdArgs a, const void* tabv, const void* twp,
                                             void* scratch) {
  __shared__ __align__(16) char lds[FwdLds<T, MODE, FUSED, V>::BYTES];
  fwd_body<T, MODE, FUSED, V>(a, xcd_block(blockIdx.x, gridDim.x), tabv, twp, scratch, lds);
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
// launchers
// ---------------------------------------------------------------------------
template <int T, int MODE, int V>
static int inv_tv(const Tables& tb, bool fused, const InvArgs& a, hipStream_t s) {
  const long nblk = inv_items<T, MODE, V>(a);
  if (nblk <= 0) return 0;
  const void *A0 = nullptr, *A1 = nullptr;
  const int t0 = (a.kind == INV_UV) ? TQV : a.table;
  if (MODE == MODE_BF16X3) {
    A0 = tb.inv_bf[t0];
    A1 = tb.inv_bf[TQU];
  } else if (MODE == MODE_F32MFMA) {
    A0 = tb.inv_f32[t0];
    A1 = tb.inv_f32[TQU];
  } else if (MODE == MODE_F64) {
    A0 = tb.inv_f64[t0];
    A1 = tb.inv_f64[TQU];
  }
  const void* tv = (MODE == MODE_F64) ? (const void*)tb.filt64 : (const void*)tb.filt32;
  const void* tw = (MODE == MODE_F64) ? (const void*)tb.tw64 : (const void*)tb.tw32;
  dim3 grid((unsigned)nblk), block(InvCfg<T, MODE, V>::NT);
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
  if (fused && a.var == 2) return inv_tv<T, MODE, 2>(tb, fused, a, s);
  if (fused && a.var == 3) return inv_tv<T, MODE, 3>(tb, fused, a, s);
  return inv_tv<T, MODE, 0>(tb, fused, a, s);
}

template <int T, int MODE, int V>
static int inv_rt_tv(const Tables& tb, const InvArgs& aq, const InvArgs& aj, const InvArgs& au,
                     hipStream_t s) {
  using C = InvCfg<T, MODE, V>;
  constexpr int NPB = TS<T>::NPAIR / C::LP;
  if (aq.nmem != aj.nmem || aq.nmem != au.nmem) return -1;
  const int gq = (aq.nfield + C::FB - 1) / C::FB;
  const int gj = (aj.nfield + C::FB - 1) / C::FB;
  const int gu = (au.nlev + C::FB / 2 - 1) / (C::FB / 2);
  const long nblk = (long)aq.nmem * NPB * (gq + gj + gu);
  if (nblk <= 0) return 0;
  const void *tq = nullptr, *tj = nullptr, *tv0 = nullptr, *tu = nullptr;
  if (MODE == MODE_BF16X3) {
    tq = tb.inv_bf[TQI];
    tj = tb.inv_bf[TQJ];
    tv0 = tb.inv_bf[TQV];
    tu = tb.inv_bf[TQU];
  } else if (MODE == MODE_F32MFMA) {
    tq = tb.inv_f32[TQI];
    tj = tb.inv_f32[TQJ];
    tv0 = tb.inv_f32[TQV];
    tu = tb.inv_f32[TQU];
  } else if (MODE == MODE_F64) {
    tq = tb.inv_f64[TQI];
    tj = tb.inv_f64[TQJ];
    tv0 = tb.inv_f64[TQV];
    tu = tb.inv_f64[TQU];
  }
  const void* tv = (MODE == MODE_F64) ? (const void*)tb.filt64 : (const void*)tb.filt32;
  const void* tw = (MODE == MODE_F64) ? (const void*)tb.tw64 : (const void*)tb.tw32;
  k_inv_rt<T, MODE, V><<<(unsigned)nblk, C::NT, 0, s>>>(aq, aj, au, tq, tj, tv0, tu, tb.inv_elems,
                                                       tv, tw, gq, gj, gu);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <int T, int MODE>
static int inv_rt_t(const Tables& tb, const InvArgs& aq, const InvArgs& aj, const InvArgs& au,
                    hipStream_t s) {
  if (aq.var == 1) return inv_rt_tv<T, MODE, 1>(tb, aq, aj, au, s);
  if (aq.var == 2) return inv_rt_tv<T, MODE, 2>(tb, aq, aj, au, s);
  if (aq.var == 3) return inv_rt_tv<T, MODE, 3>(tb, aq, aj, au, s);
  return inv_rt_tv<T, MODE, 0>(tb, aq, aj, au, s);
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
  const long nblk = fwd_items<T, MODE, V>(a);
  if (nblk <= 0) return 0;
  if (C::NPB > 1 && (scratch == nullptr || sbytes < fwd_scratch_t<T, MODE>(a.nmem, a.nfield)))
    return -4;
  // the direct sums read FP64 tables [t][pair][w], FP32 ones grouped by 4 pairs
  const void* tv = (MODE == MODE_F64) ? (const void*)tb.filt64 : (const void*)tb.fwd4;
  const void* tw = (MODE == MODE_F64) ? (const void*)tb.tw64 : (const void*)tb.tw32;
  dim3 grid((unsigned)nblk), block(C::NT);
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
  if (fused && a.var == 2) return fwd_tv<T, MODE, 2>(tb, fused, a, scratch, sbytes, s);
  if (fused && a.var == 3) return fwd_tv<T, MODE, 3>(tb, fused, a, scratch, sbytes, s);
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
int inv_mode(const Tables& tb, bool fused, const InvArgs& a, hipStream_t s) {
  EXO_DISPATCH_T(inv_t, tb, fused, a, s)
}
template <int MODE_X>
int inv_rt_mode(const Tables& tb, const InvArgs& aq, const InvArgs& aj, const InvArgs& au,
                hipStream_t s) {
  EXO_DISPATCH_T(inv_rt_t, tb, aq, aj, au, s)
}
template <int MODE_X>
int fwd_mode(const Tables& tb, bool fused, const FwdArgs& a, void* scr, size_t sb,
                    hipStream_t s) {
  EXO_DISPATCH_T(fwd_t, tb, fused, a, scr, sb, s)
}
template <int MODE_X>
size_t scr_mode(const Tables& tb, int nmem, int nfield) {
  switch (tb.ntru) {
    case 21: return fwd_scratch_t<21, MODE_X>(nmem, nfield);
    case 42: return fwd_scratch_t<42, MODE_X>(nmem, nfield);
    case 63: return fwd_scratch_t<63, MODE_X>(nmem, nfield);
    case 127: return fwd_scratch_t<127, MODE_X>(nmem, nfield);
    default: return 0;
  }
}

}  // namespace exo
