// sht_common.h -- shared device definitions of the MI355X spectral transform.
//
// Conventions (the reference's, exoplasim/plasim/src/legmod.f90 and
// fftmod.f90):
//   spectral  sp[2*w + ri], w = w0(m) + (n - m), m-major, n inner
//   Fourier   fc[(lat*(NLON/2) + m)*2 + ri], lat 0..NLAT-1 north->south
//   grid      gp[lat*NLON + j]
// Latitude pairs: pair p holds lat p (north) and NLAT-1-p (south).  The
// symmetric ("sym") class of a spectral mode is n-m even for P-type tables
// (qi, qc, qm, qq, qu) and n-m odd for Q-type tables (qj, qe, qv), since
// P(-mu) = (-1)^(n-m) P(mu) and Q(-mu) = -(-1)^(n-m) Q(mu).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace exo {

enum Mode : int { MODE_F64 = 0, MODE_F32 = 1, MODE_F32MFMA = 2, MODE_BF16X3 = 3 };
enum Table : int { TQI = 0, TQJ, TQC, TQE, TQM, TQQ, TQU, TQV, NTAB };

__host__ __device__ constexpr bool table_is_q(int t) {
  return t == TQJ || t == TQE || t == TQV;
}

// Sizes of one triangular truncation on its Gaussian grid
// (src/plasimmod.f90:100-113 with NPRO=1).
template <int T>
struct TS {
  static constexpr int NTRU = T;
  static constexpr int NM = T + 1;
  static constexpr int NLAT = (T == 21) ? 32 : (T == 42) ? 64 : (T == 63) ? 96 : 192;
  static constexpr int NLON = 2 * NLAT;
  static constexpr int NPAIR = NLAT / 2;
  static constexpr int NCSP = (T + 1) * (T + 2) / 2;
  static constexpr int NRSP = 2 * NCSP;
  // K blocks of 32 spectral rows per zonal wavenumber (inverse MFMA A operand)
  static constexpr int KB = (NM + 31) / 32;
  // max number of modes of one symmetry class per m
  static constexpr int KHALF = (NM + 1) / 2;
  // row blocks of 16 per (m, class) for the direct MFMA
  static constexpr int RB = (KHALF + 15) / 16;
};

__host__ __device__ constexpr int w0_of(int ntru, int m) {
  return m * (ntru + 1) - m * (m - 1) / 2;
}

// zonal wavenumber of spectral index w: the root of w0_of(m) = w, then one
// integer correction either way
__device__ __forceinline__ int m_of_w(int ntru, int w) {
  const int b = 2 * ntru + 3;
  int m = (int)(0.5f * ((float)b - sqrtf((float)(b * b - 8 * w))));
  m = m < 0 ? 0 : (m > ntru ? ntru : m);
  if (m < ntru && w0_of(ntru, m + 1) <= w) ++m;
  if (m > 0 && w0_of(ntru, m) > w) --m;
  return m;
}

// Operands of the split-BF16 inverse Legendre GEMM.  The reduction runs
// over total wavenumbers n in 32-wide blocks kb per zonal wavenumber m,
// each block in parity order: k = j holds n - m = 32 kb + 2 j and
// k = 16 + j holds n - m = 32 kb + 2 j + 1 (j = 0..15).  Both operands are
// stored as rows of one parity h: element j of a row is n - m = 2 j + h,
// pk_rw(m) = 8 ceil((NTRU + 1 - m) / 16) values per row, zero for n > NTRU.
// pk_c16(x) = sum_{L=1..x} 8 ceil(L/16) sums the row lengths.  Both take
// m <= NTRU + 1 and x >= 0 and compute in unsigned shifts (these offsets are
// evaluated for every work item of the inverse, in scalar registers).
__host__ __device__ constexpr int pk_rw(int ntru, int m) {
  return 8 * (int)((unsigned)(ntru + 16 - m) >> 4);
}
__host__ __device__ constexpr int pk_c16(int x) {
  const unsigned a = (unsigned)x >> 4, b = (unsigned)x & 15u;
  return (int)(64u * a * (a + 1u) + 8u * b * (a + 1u));
}
// spectral operand: per field, for each m, [re/im][h][pk_rw] fp32 values in
// parity order (the inverse splits them into bf16 planes in registers)
__host__ __device__ constexpr int pk_field(int ntru) { return 4 * pk_c16(ntru + 1); }
__host__ __device__ constexpr int pk_seg(int ntru, int m) {
  return 4 * (pk_c16(ntru + 1) - pk_c16(ntru + 1 - m));
}
// table operand: per block of 8 latitude pairs, for each m,
// [plane][row = 2 pair + class][pk_rw]; the row holds the parity of its
// class (h = class for P-type tables, 1 - class for Q-type), the other
// parity being zero and not stored
__host__ __device__ constexpr int tab_pb(int ntru) { return 48 * pk_c16(ntru + 1); }
__host__ __device__ constexpr int tab_seg(int ntru, int m) {
  return 48 * (pk_c16(ntru + 1) - pk_c16(ntru + 1 - m));
}

// Raw buffer loads: the per-lane byte offset in a VGPR and a wave-uniform
// byte offset in an SGPR, so a loop over a uniform stride (a table's
// latitude rows) costs no vector address arithmetic.  Out-of-range loads
// return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
template <typename R>
__device__ __forceinline__ R buffer_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(R) == 4)
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ uint4 buffer_load_x4(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return uint4{v[0], v[1], v[2], v[3]};
}
// x from another lane of the same group of four (DPP quad permutation,
// CTRL = p0 | p1 << 2 | p2 << 4 | p3 << 6): no LDS round trip
template <int CTRL>
__device__ __forceinline__ float quad_perm(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ unsigned quad_perm_u(unsigned x) {
  return (unsigned)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double quad_perm(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// a per-lane byte offset that no buffer covers: such a load returns 0
constexpr int BUF_NONE = 0x7ff00000;

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2: block b runs on XCD b % 8.  This renumbering gives each XCD one
// contiguous range of logical blocks, so blocks with neighbouring logical
// indices (the latitude-pair blocks of one member's field group, which read
// the same spectral coefficients) share an L2 instead of fetching the same
// data from HBM once per XCD.  A bijection on [0, nblk) whatever the
// dispatch order.
__device__ __forceinline__ int xcd_block(int b, int nblk) {
  constexpr int NX = 8;
  const int q = nblk / NX, r = nblk % NX;
  const int x = b % NX, i = b / NX;
  return x * q + (x < r ? x : r) + i;
}

template <typename R>
struct C2 {
  R x, y;
};

template <typename R>
__device__ __forceinline__ C2<R> cadd(C2<R> a, C2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ C2<R> csub(C2<R> a, C2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ C2<R> cmul(C2<R> a, C2<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by s*i (s = +1 or -1)
template <int S, typename R>
__device__ __forceinline__ C2<R> cmul_si(C2<R> a) {
  return S > 0 ? C2<R>{-a.y, a.x} : C2<R>{a.y, -a.x};
}

// ---- small in-register DFTs; S = -1 forward (e^{-2 pi i/R}), +1 inverse ----
template <int R, int S, typename F>
struct Dft;

template <int S, typename F>
struct Dft<2, S, F> {
  __device__ __forceinline__ static void run(C2<F>* v) {
    C2<F> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int S, typename F>
struct Dft<3, S, F> {
  __device__ __forceinline__ static void run(C2<F>* v) {
    const F h = F(0.86602540378443864676);  // sqrt(3)/2
    C2<F> t = cadd(v[1], v[2]);
    C2<F> d = csub(v[1], v[2]);
    C2<F> m = {v[0].x - F(0.5) * t.x, v[0].y - F(0.5) * t.y};
    C2<F> q = cmul_si<S>(C2<F>{h * d.x, h * d.y});
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, q);
    v[2] = csub(m, q);
  }
};

template <int S, typename F>
struct Dft<4, S, F> {
  __device__ __forceinline__ static void run(C2<F>* v) {
    C2<F> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C2<F> t2 = cadd(v[1], v[3]), t3 = cmul_si<S>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <int S, typename F>
struct Dft<8, S, F> {
  __device__ __forceinline__ static void run(C2<F>* v) {
    C2<F> e[4] = {v[0], v[2], v[4], v[6]};
    C2<F> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, S, F>::run(e);
    Dft<4, S, F>::run(o);
    const F r = F(0.70710678118654752440);
    // w^1 = r(1 + s i), w^2 = s i, w^3 = r(-1 + s i)
    C2<F> o1 = cmul(o[1], C2<F>{r, F(S) * r});
    C2<F> o2 = cmul_si<S>(o[2]);
    C2<F> o3 = cmul(o[3], C2<F>{-r, F(S) * r});
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// 6 = 3 x 2: DFT3 of the even and the odd inputs, a[k1] and b[k1]; then
// X[k1] = a[k1] + w6^k1 b[k1] and X[k1 + 3] = a[k1] - w6^k1 b[k1]
template <int S, typename F>
struct Dft<6, S, F> {
  __device__ __forceinline__ static void run(C2<F>* v) {
    const F h = F(0.86602540378443864676);  // sqrt(3)/2
    C2<F> a[3] = {v[0], v[2], v[4]};
    C2<F> b[3] = {v[1], v[3], v[5]};
    Dft<3, S, F>::run(a);
    Dft<3, S, F>::run(b);
    const C2<F> b1 = cmul(b[1], C2<F>{F(0.5), F(S) * h});   // w6
    const C2<F> b2 = cmul(b[2], C2<F>{F(-0.5), F(S) * h});  // w6^2
    v[0] = cadd(a[0], b[0]);
    v[3] = csub(a[0], b[0]);
    v[1] = cadd(a[1], b1);
    v[4] = csub(a[1], b1);
    v[2] = cadd(a[2], b2);
    v[5] = csub(a[2], b2);
  }
};

// 16 = 4 x 4: DFT4 over n1 of v[4 n1 + n2], twiddle w16^(n2 k1), DFT4 over
// n2 into X[k1 + 4 k2]
template <int S, typename F>
struct Dft<16, S, F> {
  __device__ __forceinline__ static void run(C2<F>* v) {
    const F c1 = F(0.92387953251128675613), s1 = F(0.38268343236508977173);
    const F r = F(0.70710678118654752440);
    C2<F> t[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      C2<F> a[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      Dft<4, S, F>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) t[n2][k1] = a[k1];
    }
    // w16^m = (cos, S sin)(2 pi m / 16) for m = n2 k1
    t[1][1] = cmul(t[1][1], C2<F>{c1, F(S) * s1});   // m = 1
    t[1][2] = cmul(t[1][2], C2<F>{r, F(S) * r});     // m = 2
    t[1][3] = cmul(t[1][3], C2<F>{s1, F(S) * c1});   // m = 3
    t[2][1] = cmul(t[2][1], C2<F>{r, F(S) * r});     // m = 2
    t[2][2] = cmul_si<S>(t[2][2]);                   // m = 4
    t[2][3] = cmul(t[2][3], C2<F>{-r, F(S) * r});    // m = 6
    t[3][1] = cmul(t[3][1], C2<F>{s1, F(S) * c1});   // m = 3
    t[3][2] = cmul(t[3][2], C2<F>{-r, F(S) * r});    // m = 6
    t[3][3] = cmul(t[3][3], C2<F>{-c1, -F(S) * s1}); // m = 9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      C2<F> b[4] = {t[0][k1], t[1][k1], t[2][k1], t[3][k1]};
      Dft<4, S, F>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
  }
};

// Radix plan of the row lengths used by the supported truncations.
template <int N>
struct FftPlan;
template <> struct FftPlan<64>  { static constexpr int NP = 2; static constexpr int R[4] = {8, 8, 1, 1}; };
template <> struct FftPlan<128> { static constexpr int NP = 2; static constexpr int R[4] = {16, 8, 1, 1}; };
template <> struct FftPlan<192> { static constexpr int NP = 3; static constexpr int R[4] = {8, 8, 3, 1}; };
template <> struct FftPlan<384> { static constexpr int NP = 3; static constexpr int R[4] = {16, 8, 3, 1}; };

// FFTs in flight per pass within an NT-thread block (about 8 points per
// thread)
template <int N, int NT = 256>
struct FftBatch {
  // N = 192 on 512 threads: 16 rows (three batches for the 48 rows of a T63
  // direct block), so that the block's LDS (77 KB) lets two blocks share a CU
  static constexpr int NF =
      ((N == 64) ? 32 : (N == 128) ? 16 : (N == 192) ? ((NT == 512) ? 8 : 10) : 5) * (NT / 256);
};

// FFT work rows in LDS hold one pad element after every 32 complex values:
// the first pass writes with a stride of 8 or 16 complex (64 or 128 bytes),
// which without the pad would put 16 or 32 lanes on each bank.
__host__ __device__ constexpr int fft_buf_elems(int n) { return n + (n >> 5); }
__device__ __forceinline__ int bix(int i) { return i + (i >> 5); }

// The thread index as an opaque value.  A persistent kernel runs many work
// items of different shapes in one loop; without this the compiler hoists
// each item type's thread-index arithmetic out of that loop and keeps all of
// it in registers at once.
__device__ __forceinline__ int item_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// fn(integral_constant<int, I>) for I = B .. E-1
template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  if constexpr (B < E) {
    fn(std::integral_constant<int, B>{});
    static_for<B + 1, E>(fn);
  }
}

// Padded index of in-row element base(j) + K over the butterflies j of a
// pass (base(j) = j for the reads, the Stockham output index for the
// writes): if the pad count of base(j) + K exceeds that of base(j) by the
// same D for every j, the padded index is that of base(j) plus K + D, a
// constant offset of one per-butterfly address; -1 otherwise.
template <int NB, int NS, int R, int K, bool OUT>
constexpr int pad_delta() {
  int d = 0;
  for (int j = 0; j < NB; ++j) {
    const int base = OUT ? (j / NS) * NS * R + j % NS : j;
    const int dd = ((base + K) >> 5) - (base >> 5);
    if (j == 0)
      d = dd;
    else if (dd != d)
      return -1;
  }
  return d;
}

// One Stockham autosort pass (radix R, stride NS) over nf (<= NFM) FFTs of
// length N (a multiple of 32) held in LDS rows of N + N/32 complex.  With
// FIRST the pass reads through `load(f, k)`, else from `buf`; with LAST it
// writes through `store(f, k, v)`, else to `buf`.  The first pass (NS = 1)
// of a transform read from `buf` treats the inputs ZLO..ZHI as zero without
// reading them when a butterfly's whole range lies there.  All NT threads
// of the block must call it.  tw[k] = e^{S 2 pi i k/N}.
template <int N, int R, int NS, int S, bool FIRST, bool LAST, typename F, int NT, int NFM,
          int ZLO, int ZHI, typename Load, typename Store>
__device__ __forceinline__ void fft_pass(C2<F>* buf, const C2<F>* tw, int nf,
                                         Load& load, Store& store) {
  static_assert(N % 32 == 0, "rows are padded per 32 complex");
  constexpr int NB = N / R;  // butterflies per FFT
  constexpr int NP = N + N / 32;
  constexpr int BPT = (NFM * NB + NT - 1) / NT;
  const int tid = item_tid();
  __builtin_assume(tid >= 0 && tid < NT);
  const int total = nf * NB;
  C2<F> v[BPT][R];
#pragma unroll
  for (int t = 0; t < BPT; ++t) {
    const int b = tid + t * NT;
    if (b < total) {
      const int f = b / NB, j = b - f * NB;
      const int rb = f * NP + j + (j >> 5);
      static_for<0, R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        constexpr int K = r * NB;
        constexpr int D = pad_delta<NB, NS, R, K, false>();
        if constexpr (FIRST) {
          v[t][r] = load(f, j + K);
        } else if constexpr (NS == 1 && K >= ZLO && K + NB - 1 <= ZHI) {
          v[t][r] = C2<F>{F(0), F(0)};
        } else if constexpr (D >= 0) {
          v[t][r] = buf[rb + K + D];
        } else {
          v[t][r] = buf[f * NP + j + K + ((j + K) >> 5)];
        }
      });
    }
  }
  // in place (the last pass's store may write the same rows too): every
  // read of the pass ends before the first write
  if (!FIRST) __syncthreads();
#pragma unroll
  for (int t = 0; t < BPT; ++t) {
    const int b = tid + t * NT;
    if (b < total) {
      const int f = b / NB, j = b - f * NB;
      const int jm = j % NS;
      if (NS > 1) {
        // jm r N/(NS R) < N
#pragma unroll
        for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], tw[jm * r * (N / (NS * R))]);
      }
      Dft<R, S, F>::run(v[t]);
      const int idxd = (j / NS) * NS * R + jm;
      const int wb = f * NP + idxd + (idxd >> 5);
      static_for<0, R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        constexpr int K = r * NS;
        constexpr int D = pad_delta<NB, NS, R, K, true>();
        if constexpr (LAST)
          store(f, idxd + K, v[t][r]);
        else if constexpr (D >= 0)
          buf[wb + K + D] = v[t][r];
        else
          buf[f * NP + idxd + K + ((idxd + K) >> 5)] = v[t][r];
      });
    }
  }
  __syncthreads();
}

// Full FFT of nf (<= FftBatch<N, NT>::NF) sequences through the radix plan
// by a block of NT threads: the first pass reads through `load`, the last
// writes through `store`.
template <int N, int S, typename F, int NT = 256, typename Load, typename Store>
__device__ __forceinline__ void fft_run(C2<F>* buf, const C2<F>* tw, int nf, Load& load,
                                        Store& store) {
  using P = FftPlan<N>;
  constexpr int R0 = P::R[0], R1 = P::R[1], R2 = P::R[2], R3 = P::R[3];
  constexpr int M = FftBatch<N, NT>::NF;
  if constexpr (P::NP == 2) {
    fft_pass<N, R0, 1, S, true, false, F, NT, M, 1, 0>(buf, tw, nf, load, store);
    fft_pass<N, R1, R0, S, false, true, F, NT, M, 1, 0>(buf, tw, nf, load, store);
  } else if constexpr (P::NP == 3) {
    fft_pass<N, R0, 1, S, true, false, F, NT, M, 1, 0>(buf, tw, nf, load, store);
    fft_pass<N, R1, R0, S, false, false, F, NT, M, 1, 0>(buf, tw, nf, load, store);
    fft_pass<N, R2, R0 * R1, S, false, true, F, NT, M, 1, 0>(buf, tw, nf, load, store);
  } else {
    fft_pass<N, R0, 1, S, true, false, F, NT, M, 1, 0>(buf, tw, nf, load, store);
    fft_pass<N, R1, R0, S, false, false, F, NT, M, 1, 0>(buf, tw, nf, load, store);
    fft_pass<N, R2, R0 * R1, S, false, false, F, NT, M, 1, 0>(buf, tw, nf, load, store);
    fft_pass<N, R3, R0 * R1 * R2, S, false, true, F, NT, M, 1, 0>(buf, tw, nf, load, store);
  }
}

// The first pass (radix R = FftPlan<N>::R[0], stride 1) of a Stockham FFT
// whose inputs the calling thread already holds: butterfly b = f NB + j
// (NB = N / R) takes inputs k = j + r NB, r < R, from v.  Threads with
// active == false hold no butterfly of this batch.  The outputs go to the
// rows of `buf` (N + N/32 complex each); all threads of the block must call
// it.
template <int N, int S, typename F>
__device__ __forceinline__ void fft_pass0_held(C2<F>* buf, bool active, int b,
                                               const C2<F> (&v)[FftPlan<N>::R[0]]) {
  constexpr int R = FftPlan<N>::R[0];
  constexpr int NB = N / R;
  constexpr int NP = N + N / 32;
  if (active) {
    C2<F> w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = v[r];
    Dft<R, S, F>::run(w);
    const int f = b / NB, j = b - f * NB;
    const int idxd = j * R;
    const int wb = f * NP + idxd + (idxd >> 5);
    static_for<0, R>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      constexpr int D = pad_delta<NB, 1, R, r, true>();
      if constexpr (D >= 0)
        buf[wb + r + D] = w[r];
      else
        buf[f * NP + idxd + r + ((idxd + r) >> 5)] = w[r];
    });
  }
  __syncthreads();
}

// The passes after the first of the FFTs of nf (<= NFM) sequences whose
// first pass is in `buf`; the last pass writes through `store`.
template <int N, int S, typename F, int NT, int NFM, typename Store>
__device__ __forceinline__ void fft_run_after0(C2<F>* buf, const C2<F>* tw, Store& store,
                                               int nf) {
  using P = FftPlan<N>;
  constexpr int R0 = P::R[0], R1 = P::R[1], R2 = P::R[2], R3 = P::R[3];
  auto none = [](int, int) { return C2<F>{F(0), F(0)}; };
  if constexpr (P::NP == 2) {
    fft_pass<N, R1, R0, S, false, true, F, NT, NFM, 1, 0>(buf, tw, nf, none, store);
  } else if constexpr (P::NP == 3) {
    fft_pass<N, R1, R0, S, false, false, F, NT, NFM, 1, 0>(buf, tw, nf, none, store);
    fft_pass<N, R2, R0 * R1, S, false, true, F, NT, NFM, 1, 0>(buf, tw, nf, none, store);
  } else {
    fft_pass<N, R1, R0, S, false, false, F, NT, NFM, 1, 0>(buf, tw, nf, none, store);
    fft_pass<N, R2, R0 * R1, S, false, false, F, NT, NFM, 1, 0>(buf, tw, nf, none, store);
    fft_pass<N, R3, R0 * R1 * R2, S, false, true, F, NT, NFM, 1, 0>(buf, tw, nf, none, store);
  }
}

// Full FFT of nf (<= NFM) sequences already in `buf` (rows of N + N/32
// complex), in place; with OUT the last pass writes through `store`
// instead.  Inputs ZLO..ZHI of every row are zero (and must be zero in
// `buf`).
template <int N, int S, typename F, int NT, int NFM, int ZLO, int ZHI, bool OUT = true,
          typename Store>
__device__ __forceinline__ void fft_run_lds(C2<F>* buf, const C2<F>* tw, Store& store,
                                            int nf = NFM) {
  using P = FftPlan<N>;
  constexpr int R0 = P::R[0], R1 = P::R[1], R2 = P::R[2], R3 = P::R[3];
  auto none = [](int, int) { return C2<F>{F(0), F(0)}; };
  if constexpr (P::NP == 2) {
    fft_pass<N, R0, 1, S, false, false, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none, store);
    fft_pass<N, R1, R0, S, false, OUT, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none, store);
  } else if constexpr (P::NP == 3) {
    fft_pass<N, R0, 1, S, false, false, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none, store);
    fft_pass<N, R1, R0, S, false, false, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none, store);
    fft_pass<N, R2, R0 * R1, S, false, OUT, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none, store);
  } else {
    fft_pass<N, R0, 1, S, false, false, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none, store);
    fft_pass<N, R1, R0, S, false, false, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none, store);
    fft_pass<N, R2, R0 * R1, S, false, false, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none, store);
    fft_pass<N, R3, R0 * R1 * R2, S, false, OUT, F, NT, NFM, ZLO, ZHI>(buf, tw, nf, none,
                                                                      store);
  }
}

// The LDS writes of one wave before this point are seen by the LDS reads of
// its other lanes after it: a wave's LDS instructions execute in order, so
// only the compiler must keep them in program order.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FFT of one row of N = 64 R0 complex values (R0 = 3 or 6: 192 or 384) by
// one wave (lane = 0..63), with passes of radix R0, 8 and 8 (Stockham
// autosort; the first pass has one butterfly per lane).  The first pass
// reads input k through load(r, k), k = lane + 64 r; the transform is left
// in `row` (N + N/32 complex, padded as bix).  tw[k] = e^{S 2 pi i k / N}.
// Only the calling wave touches `row`.
template <int N, int S, typename F, typename Load>
__device__ __forceinline__ void fft_wave(C2<F>* row, const C2<F>* tw, int lane, Load& load) {
  constexpr int R0 = N / 64;  // first radix
  constexpr int NB = N / 8;   // butterflies of the radix-8 passes (lanes 0 .. NB-1)
  static_assert(N == 64 * R0 && (R0 == 3 || R0 == 6), "rows of 192 or 384");
  {  // radix R0, stride 1: butterfly j = lane takes k = j + 64 r
    C2<F> v[R0];
#pragma unroll
    for (int r = 0; r < R0; ++r) v[r] = load(r, lane + 64 * r);
    Dft<R0, S, F>::run(v);
#pragma unroll
    for (int r = 0; r < R0; ++r) row[bix(R0 * lane + r)] = v[r];
  }
  wave_lds_sync();
  // radix 8, stride R0, then radix 8, stride 8 R0: butterflies j < NB take
  // k = j + NB r; in place, all reads of the wave before its writes
  C2<F> v[8];
  if (lane < NB) {
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = row[bix(lane + NB * r)];
  }
  wave_lds_sync();
  if (lane < NB) {
    const int jm = lane % R0;
#pragma unroll
    for (int r = 1; r < 8; ++r) v[r] = cmul(v[r], tw[jm * r * 8]);
    Dft<8, S, F>::run(v);
    const int o = (lane / R0) * (8 * R0) + jm;
#pragma unroll
    for (int r = 0; r < 8; ++r) row[bix(o + R0 * r)] = v[r];
  }
  wave_lds_sync();
  if (lane < NB) {
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = row[bix(lane + NB * r)];
  }
  wave_lds_sync();
  if (lane < NB) {
#pragma unroll
    for (int r = 1; r < 8; ++r) v[r] = cmul(v[r], tw[lane * r]);
    Dft<8, S, F>::run(v);
#pragma unroll
    for (int r = 0; r < 8; ++r) row[bix(lane + NB * r)] = v[r];
  }
  wave_lds_sync();
}

// ---- split-BF16 helpers ----
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf16_trunc_bits(float x) {
  return (uint16_t)(__float_as_uint(x) >> 16);
}

// exact three-way split x = hi + mid + lo of an FP32 value into bf16
// (truncation at each step, so each remainder is exact)
__device__ __forceinline__ void split3(float x, uint16_t& h, uint16_t& m, uint16_t& l) {
  uint32_t u = __float_as_uint(x);
  float fh = __uint_as_float(u & 0xffff0000u);
  float r1 = x - fh;
  uint32_t u1 = __float_as_uint(r1);
  float fm = __uint_as_float(u1 & 0xffff0000u);
  float r2 = r1 - fm;
  h = (uint16_t)(u >> 16);
  m = (uint16_t)(u1 >> 16);
  l = (uint16_t)(__float_as_uint(r2) >> 16);
}

union BF8 {
  bf16x8 v;
  uint16_t s[8];
  uint4 q;
};

}  // namespace exo
