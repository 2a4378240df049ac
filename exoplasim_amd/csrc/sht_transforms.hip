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
//   MODE_F64      inverse: v_mfma_f64_16x16x4_f64; direct: FP64 FMA on the
//                 vector ALUs (the plain path)
//   MODE_F32      FP32 FMA on the vector ALUs
//   MODE_F32MFMA  v_mfma_f32_16x16x4_f32, FP32 accumulate
//   MODE_BF16X3   v_mfma_f32_16x16x32_bf16 on an exact three-way bf16 split
//                 of both operands, six products (hh, hm, mh, hl, mm, lh),
//                 FP32 accumulate
// FFTs run in FP64 for MODE_F64 and FP32 otherwise.
#include "sht_bodies.h"
#include "sht_launch.h"

namespace exo {

// ---------------------------------------------------------------------------
// FFT of independent rows (in place), two real rows per complex FFT
// DIR +1: Fourier -> grid (fc2gp), DIR -1: grid -> Fourier (gp2fc)
// ---------------------------------------------------------------------------
// blockIdx.y selects a member: its rows start mstride values after the
// previous member's.  DIR -1 reads the rows from src (smstride per member;
// src = a for the in-place transform) and writes the kmax first complex
// coefficients of each row (kmax = N/2: all of them).
template <int N, typename R, int DIR>
__global__ __launch_bounds__(256) void k_fft_rows(R* a, long lot, const C2<R>* tw,
                                                  long mstride, const R* src, long smstride,
                                                  int kmax) {
  a += blockIdx.y * mstride;
  src += blockIdx.y * smstride;
  constexpr int NF = FftBatch<N>::NF;
  __shared__ C2<R> fbuf[fft_buf_elems(NF * N)];
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
    auto srow = [&](int f, int h) -> const R* {
      const long r = r0 + 2 * f + h;
      return r < lot ? src + r * N : nullptr;
    };
    auto load = [&](int f, int k) -> C2<R> {
      const R* x0 = srow(f, 0);
      const R* x1 = srow(f, 1);
      return C2<R>{x0 ? x0[k] : R(0), x1 ? x1[k] : R(0)};
    };
    auto store = [&](int f, int k, C2<R> v) { fbuf[bix(f * N + k)] = v; };
    fft_run<N, -1, R>(fbuf, tws, nf, load, store);
    const R sc = R(0.5) / R(N);
    for (int idx = threadIdx.x; idx < nf * kmax; idx += 256) {
      const int f = idx / kmax, k = idx % kmax;
      const C2<R> zk = fbuf[bix(f * N + k)];
      const C2<R> zn = fbuf[bix(f * N + ((N - k) % N))];
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

int launch_inverse(const Tables& tb, int mode, bool fused, const InvArgs& a, hipStream_t s) {
  switch (mode) {
    case MODE_F64: return inv_mode<MODE_F64>(tb, fused, a, s);
    case MODE_F32: return inv_mode<MODE_F32>(tb, fused, a, s);
    case MODE_F32MFMA: return inv_mode<MODE_F32MFMA>(tb, fused, a, s);
    case MODE_BF16X3: return inv_mode<MODE_BF16X3>(tb, fused, a, s);
    default: return -1;
  }
}

int launch_inverse_rt(const Tables& tb, int mode, const InvArgs& aq, const InvArgs& aj,
                      const InvArgs& au, hipStream_t s) {
  switch (mode) {
    case MODE_F64: return inv_rt_mode<MODE_F64>(tb, aq, aj, au, s);
    case MODE_F32: return inv_rt_mode<MODE_F32>(tb, aq, aj, au, s);
    case MODE_F32MFMA: return inv_rt_mode<MODE_F32MFMA>(tb, aq, aj, au, s);
    case MODE_BF16X3: return inv_rt_mode<MODE_BF16X3>(tb, aq, aj, au, s);
    default: return -1;
  }
}

int launch_forward(const Tables& tb, int mode, bool fused, const FwdArgs& a, void* scratch,
                   size_t sbytes, hipStream_t s) {
  if (mode != MODE_F64 && tb.ntru == 127) {
    // T127 in the FP32 modes: the MFMA direct sums with the one table qc,
    // each field taking the place of a level; grid input first goes through
    // the row FFT into scratch (coefficients m <= T only).  The fused
    // k_fwd<127> holds one field-level per block and writes three
    // latitude-partial sums (564 us against 431 us for 121 fields x 8
    // members); at T21-T63 k_fwd is as fast or faster.
    const long fsz = (long)tb.nlat * tb.nlon;
    const void* fin = (const float*)a.in + a.src0 * fsz;
    long fstride = a.in_stride;
    if (fused) {
      if (!scratch || sbytes < fwd_fft_scratch_bytes(tb, a.nmem, a.nfield)) return -4;
      const long ms = (long)a.nfield * fsz;
      const int rc = launch_fft(tb, mode, -1, scratch, tb.nlon, (long)a.nfield * tb.nlat, s,
                                a.nmem, ms, fin, a.in_stride, tb.ntru + 1);
      if (rc) return rc;
      fin = scratch;
      fstride = ms;
    }
    DirxArgs d{};
    d.in[0] = fin;
    d.out = (float*)a.sp + (long)a.dst0 * 2 * tb.ncsp;
    d.nin = 1;
    d.nlev = a.nfield;
    d.nmem = a.nmem;
    d.in_stride = fstride;
    d.out_stride = a.sp_stride;
    d.nterm = 2;
    d.terms[0] = DirxTerm{0, 0, 0, +1, (int8_t)a.table};
    d.terms[1] = DirxTerm{0, 1, 1, +1, (int8_t)a.table};
    return launch_dirx(tb, MODE_F32, d, nullptr, 0, s);
  }
  switch (mode) {
    case MODE_F64: return fwd_mode<MODE_F64>(tb, fused, a, scratch, sbytes, s);
    case MODE_F32: return fwd_mode<MODE_F32>(tb, fused, a, scratch, sbytes, s);
    case MODE_F32MFMA: return fwd_mode<MODE_F32MFMA>(tb, fused, a, scratch, sbytes, s);
    case MODE_BF16X3: return fwd_mode<MODE_BF16X3>(tb, fused, a, scratch, sbytes, s);
    default: return -1;
  }
}

size_t fwd_fft_scratch_bytes(const Tables& tb, int nmem, int nfield) {
  return sizeof(float) * (size_t)nmem * nfield * tb.nlat * tb.nlon;
}

size_t forward_scratch_bytes(const Tables& tb, int mode, int nmem, int nfield) {
  if (mode != MODE_F64 && tb.ntru == 127) return fwd_fft_scratch_bytes(tb, nmem, nfield);
  switch (mode) {
    case MODE_F64: return scr_mode<MODE_F64>(tb, nmem, nfield);
    case MODE_F32: return scr_mode<MODE_F32>(tb, nmem, nfield);
    case MODE_F32MFMA: return scr_mode<MODE_F32MFMA>(tb, nmem, nfield);
    case MODE_BF16X3: return scr_mode<MODE_BF16X3>(tb, nmem, nfield);
    default: return 0;
  }
}

template <int N, typename R>
static int fft_t(const void* tw, int dir, R* a, long lot, int nmem, long mstride, hipStream_t s,
                 const R* src, long smstride, int kmax) {
  constexpr int NF = FftBatch<N>::NF;
  const long nblk = (lot + 2 * NF - 1) / (2 * NF);
  if (nblk <= 0 || nmem <= 0) return 0;
  const dim3 grid((unsigned)nblk, (unsigned)nmem);
  if (dir > 0)
    k_fft_rows<N, R, +1><<<grid, 256, 0, s>>>(a, lot, (const C2<R>*)tw, mstride, a, mstride,
                                               N / 2);
  else
    k_fft_rows<N, R, -1><<<grid, 256, 0, s>>>(a, lot, (const C2<R>*)tw, mstride,
                                               src ? src : a, src ? smstride : mstride,
                                               kmax > 0 && kmax < N / 2 ? kmax : N / 2);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// ---------------------------------------------------------------------------
// parity-ordered spectral operand of the split-BF16 inverse (layout in
// sht_common.h): one thread per 4 consecutive coefficients n - m = 4g ..
// 4g + 3 of one (member, field, m), read as two contiguous 16-byte loads
// (consecutive threads read consecutive runs) and written as 8 bytes, the
// values j = 2g, 2g + 1, into each of the m's four rows (re/im x parity),
// so a wave writes 512 contiguous bytes of each row.  Coefficients past the
// truncation are zeros.  (Each thread reading 8 values of one row at a
// 16-byte stride put a wave's loads on 64 cache lines each.)
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(256) void k_pack_sp(const float* sp, long sp_stride, int f0, int nf,
                                                 int nmem, float* out) {
  using S = TS<T>;
  constexpr int NC = pk_c16(T + 1) / 2;  // threads per field: sum over m of rw(m) / 2
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)nmem * nf * NC) return;
  const int u = (int)(i % NC);
  const long fm = i / NC;
  const int field = (int)(fm % nf), mem = (int)(fm / nf);
  // the largest m whose segment starts at or before thread u (segment of m:
  // threads pk_seg(m) / 8 .. + rw(m) / 2)
  int lo = 0, hi = T;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pk_seg(T, mid) / 8 <= u)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int m = lo, rw = pk_rw(T, m);
  const int g = u - pk_seg(T, m) / 8;  // coefficients n - m = 4g .. 4g + 3
  // the member's fields as one buffer: a run past its end reads zeros, one
  // past the truncation (into the next m or field) is zeroed below
  const float* base = sp + mem * sp_stride;
  const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(base, (int)((f0 + nf) * (long)S::NRSP * 4));
  const int off = (int)(((long)(f0 + field) * S::NRSP + 2 * (w0_of(T, m) + 4 * g)) * 4);
  const uint4 a0 = buffer_load_x4(rs, off, 0), a1 = buffer_load_x4(rs, off, 16);
  float x[8] = {__uint_as_float(a0.x), __uint_as_float(a0.y), __uint_as_float(a0.z),
                __uint_as_float(a0.w), __uint_as_float(a1.x), __uint_as_float(a1.y),
                __uint_as_float(a1.z), __uint_as_float(a1.w)};
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (m + 4 * g + q > T) x[2 * q] = x[2 * q + 1] = 0.f;
  // rows (re/im ri, parity h) = 2 ri + h; values j = 2g (n - m = 4g + h) and
  // 2g + 1 (4g + 2 + h)
  float* dst = out + fm * pk_field(T) + pk_seg(T, m) + 2 * g;
#pragma unroll
  for (int ri = 0; ri < 2; ++ri)
#pragma unroll
    for (int h = 0; h < 2; ++h)
      *(float2*)(dst + (2 * ri + h) * rw) = make_float2(x[2 * h + ri], x[2 * (2 + h) + ri]);
}

template <int T>
static int pack_t(const float* sp, long sps, int f0, int nf, int nmem, float* out,
                  hipStream_t s) {
  const long total = (long)nmem * nf * (pk_c16(T + 1) / 2);
  if (total <= 0) return 0;
  k_pack_sp<T><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(sp, sps, f0, nf, nmem, out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_pack(const Tables& tb, const float* sp, long sp_stride, int f0, int nf, int nmem,
                float* out, hipStream_t s) {
  switch (tb.ntru) {
    case 21: return pack_t<21>(sp, sp_stride, f0, nf, nmem, out, s);
    case 42: return pack_t<42>(sp, sp_stride, f0, nf, nmem, out, s);
    case 63: return pack_t<63>(sp, sp_stride, f0, nf, nmem, out, s);
    case 127: return pack_t<127>(sp, sp_stride, f0, nf, nmem, out, s);
    default: return -2;
  }
}

int launch_fft(const Tables& tb, int mode, int dir, void* a, int n, long lot, hipStream_t s,
               int nmem, long mstride, const void* src, long smstride, int kmax) {
  if (n != tb.nlon) return -2;
  const bool d = (mode == MODE_F64);
#define EXO_FFT_CASE(NN)                                                                    \
  case NN:                                                                                  \
    return d ? fft_t<NN, double>(tb.tw64, dir, (double*)a, lot, nmem, mstride, s,           \
                                 (const double*)src, smstride, kmax)                        \
             : fft_t<NN, float>(tb.tw32, dir, (float*)a, lot, nmem, mstride, s,             \
                                (const float*)src, smstride, kmax);
  switch (n) {
    EXO_FFT_CASE(64)
    EXO_FFT_CASE(128)
    EXO_FFT_CASE(192)
    EXO_FFT_CASE(384)
    default: return -2;
  }
#undef EXO_FFT_CASE
}

}  // namespace exo
