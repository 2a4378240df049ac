// rad_mlp.hip -- forward and backward kernels of the long-wave emulator MLP
// (43 -> 256 -> 256 -> 256 -> 21, GELU), BF16 operands on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Layouts (row-major, leading dimension = the padded feature count):
//   activations X[M][K] bf16, weights W[N][K] bf16 (out x in) and their
//   transpose Wt[K][N] bf16, pre-activations Z[M][N] fp32.
// K and N are multiples of 32; M (the batch) is any size.
//
//   k_fwd     Z = X W^T + b; H = gelu(Z) (bf16) or Y = Z (fp32, last layer)
//   k_bwd_dx  dH = dZ W; dZprev = dH * gelu'(Zprev) (bf16 + fp32 copies)
//   k_bwd_dw  split-K partial sums of dW = dZ^T X over the batch
//   k_sum_dw  sum of the partials; k_bias_grad partials of db = sum_m dZ
//   k_cast_w  FP32 master weights -> W, Wt bf16 (zero padded)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/exo_rad.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

union V8 {
  bf16x8 v;
  uint16_t s[8];
  uint4 q;
};

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// GELU (erf form) and its derivative.  erf by Abramowitz & Stegun 7.1.26:
// erf(x) = 1 - (a1 t + ... + a5 t^5) exp(-x^2), t = 1 / (1 + p x) for x >= 0,
// absolute error <= 1.5e-7 -- below the bf16 rounding that follows every
// activation by four orders of magnitude, and a third of the cost of the
// correctly rounded erff.  The derivative shares exp(-z^2/2).
__device__ __forceinline__ float erf_as(float x, float& e) {
  const float ax = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.0f));
  const float poly =
      t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f),
                       -0.284496736f),
               0.254829592f);
  e = __expf(-ax * ax);
  const float r = 1.0f - poly * e;
  return copysignf(r, x);
}
__device__ __forceinline__ float gelu(float z) {
  float e;
  return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f, e));
}
__device__ __forceinline__ float gelu_d(float z) {
  float e;  // exp(-z^2 / 2)
  const float cdf = 0.5f * (1.0f + erf_as(z * 0.70710678118654752f, e));
  return cdf + z * 0.39894228040143268f * e;
}

__device__ __forceinline__ f32x4 mfma(const V8& a, const V8& b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
}

// 64 x 64 output tile per 256-thread block, 32 x 32 per wave; A[M][K] and
// B as Bt[NB][K] (both K-contiguous), C[m][n] = sum_k A[m][k] Bt[n][k].
// Rows of A beyond M and of Bt beyond NB read row 0 (valid memory); their
// results are not stored.
template <typename Epi>
__device__ __forceinline__ void tile_gemm(const uint16_t* A, const uint16_t* Bt, int M, int K,
                                          int NB, int m0, int n0, Epi epi) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane >> 4, c = lane & 15;
  const int wm = m0 + (wv >> 1) * 32, wn = n0 + (wv & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  int ra[2], rb[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int r = wm + i * 16 + c;
    ra[i] = r < M ? r : 0;
    const int rn = wn + i * 16 + c;
    rb[i] = rn < NB ? rn : 0;
  }
  for (int k0 = 0; k0 < K; k0 += 32) {
    V8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i].q = *(const uint4*)(A + (long)ra[i] * K + k0 + q * 8);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j].q = *(const uint4*)(Bt + (long)rb[j] * K + k0 + q * 8);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
  }
  // C[row = 4q + r][col = c] of each 16 x 16 tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = wm + i * 16 + 4 * q + r;
        const int n = wn + j * 16 + c;
        if (m < M) epi(m, n, acc[i][j][r]);
      }
}

// forward layer: Z = X W^T + b, H = gelu(Z) (bf16) or Y = Z (fp32)
__global__ __launch_bounds__(256) void k_fwd(const uint16_t* X, const uint16_t* W, const float* b,
                                             int M, int K, int N, int nvalid, int act, float* Z,
                                             uint16_t* H, float* Y) {
  const int n0 = (blockIdx.x % (N / 64 > 0 ? N / 64 : 1)) * 64;
  const int m0 = (blockIdx.x / (N / 64 > 0 ? N / 64 : 1)) * 64;
  tile_gemm(X, W, M, K, N, m0, n0, [&](int m, int n, float v) {
    if (n >= N) return;
    const float z = v + (n < nvalid ? b[n] : 0.f);
    if (act) {
      if (Z) Z[(long)m * N + n] = z;
      H[(long)m * N + n] = (n < nvalid) ? f2bf(gelu(z)) : (uint16_t)0;
    } else {
      Y[(long)m * N + n] = z;
    }
  });
}

// dH = dZ W (dZ[M][N], W as Wt[K][N]); dZprev = dH * gelu'(Zprev)
__global__ __launch_bounds__(256) void k_bwd_dx(const uint16_t* dZ, const uint16_t* Wt,
                                                const float* Zprev, int M, int N, int K,
                                                int kvalid, uint16_t* dZp, float* dZp32) {
  const int nt = K / 64 > 0 ? K / 64 : 1;
  const int k0 = (blockIdx.x % nt) * 64;
  const int m0 = (blockIdx.x / nt) * 64;
  tile_gemm(dZ, Wt, M, N, K, m0, k0, [&](int m, int k, float v) {
    if (k >= K) return;
    const float g = (k < kvalid) ? v * gelu_d(Zprev[(long)m * K + k]) : 0.f;
    dZp[(long)m * K + k] = f2bf(g);
    if (dZp32) dZp32[(long)m * K + k] = g;
  });
}

// partial dW[N][K] over rows [m_begin, m_end) of the batch:
// dW[n][k] = sum_m dZ[m][n] X[m][k]; the 32 x 64 slabs of dZ and X are
// staged in LDS transposed so that the MFMA operands are m-contiguous
__global__ __launch_bounds__(256) void k_bwd_dw(const uint16_t* dZ, const uint16_t* X, int M,
                                                int N, int K, int mchunk, float* part) {
  __shared__ uint16_t sz[64][40];  // [n][m], padded rows
  __shared__ uint16_t sx[64][40];  // [k][m]
  const int ntn = N / 64 > 0 ? N / 64 : 1, ntk = K / 64 > 0 ? K / 64 : 1;
  const int tile = blockIdx.x % (ntn * ntk);
  const int split = blockIdx.x / (ntn * ntk);
  const int n0 = (tile / ntk) * 64, k0 = (tile % ntk) * 64;
  const int mb = split * mchunk, me = min(M, mb + mchunk);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane >> 4, c = lane & 15;
  const int wn = (wv >> 1) * 32, wk = (wv & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int m0 = mb; m0 < me; m0 += 32) {
    // 32 rows x 64 columns of each operand: thread t loads 8 columns of a row
    {
      const int t = threadIdx.x;
      const int r = t >> 3, cc = (t & 7) * 8;  // r in 0..31
      const int m = m0 + r;
      V8 vz, vx;
      if (m < me && n0 + cc < N) vz.q = *(const uint4*)(dZ + (long)m * N + n0 + cc);
      else vz.q = uint4{0, 0, 0, 0};
      if (m < me && k0 + cc < K) vx.q = *(const uint4*)(X + (long)m * K + k0 + cc);
      else vx.q = uint4{0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        sz[cc + j][r] = vz.s[j];
        sx[cc + j][r] = vx.s[j];
      }
    }
    __syncthreads();
    V8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) a[i].s[j] = sz[wn + i * 16 + c][q * 8 + j];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) b[i].s[j] = sx[wk + i * 16 + c][q * 8 + j];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    __syncthreads();
  }
  float* out = part + (long)split * N * K;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn + i * 16 + 4 * q + r;
        const int k = k0 + wk + j * 16 + c;
        if (n < N && k < K) out[(long)n * K + k] = acc[i][j][r];
      }
}

__global__ void k_sum_dw(const float* part, int nsplit, long nk, float* dW) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nk) return;
  float s = 0.f;
  for (int p = 0; p < nsplit; ++p) s += part[(long)p * nk + i];
  dW[i] = s;
}

// partial bias gradients: out[s][n] = sum of dZ[m][n] (fp32 input) over the
// rows m of chunk s; one block per (64 columns, chunk), summed in a fixed
// order by k_sum_dw
__global__ __launch_bounds__(256) void k_bias_grad(const float* dZ, int M, int N, int mchunk,
                                                   float* out) {
  __shared__ float red[4][64];
  const int ntn = N / 64 > 0 ? N / 64 : 1;
  const int tn = blockIdx.x % ntn, sp = blockIdx.x / ntn;
  const int n = tn * 64 + (threadIdx.x & 63);
  const int part = threadIdx.x >> 6;
  const int m1 = min(M, (sp + 1) * mchunk);
  float s = 0.f;
  if (n < N)
    for (int m = sp * mchunk + part; m < m1; m += 4) s += dZ[(long)m * N + n];
  red[part][threadIdx.x & 63] = s;
  __syncthreads();
  if (part == 0 && n < N)
    out[(long)sp * N + n] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                            red[3][threadIdx.x];
}

// FP32 master W[nout][nin] -> padded bf16 W[N][K] and Wt[K][N]
__global__ void k_cast_w(const float* w, int nout, int nin, int N, int K, uint16_t* Wb,
                         uint16_t* Wtb) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)N * K) return;
  const int n = (int)(i / K), k = (int)(i % K);
  const uint16_t v = (n < nout && k < nin) ? f2bf(w[(long)n * nin + k]) : (uint16_t)0;
  Wb[(long)n * K + k] = v;
  Wtb[(long)k * N + n] = v;
}

// fp32 -> bf16 rows with zero padding: src[M][nin] -> dst[M][K]
__global__ void k_cast_rows(const float* src, int M, int nin, int K, uint16_t* dst) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)M * K) return;
  const long m = i / K;
  const int k = (int)(i % K);
  dst[i] = k < nin ? f2bf(src[m * nin + k]) : (uint16_t)0;
}

inline unsigned nb(long n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// fused inference: features -> 4 layers -> heating rates and OLR, one block
// of 8 waves per 32 columns, activations in LDS, weights streamed from L2.
// Each layer is computed transposed, C^T[n][m] = W[n][:] . H[m][:], so that a
// lane's four accumulator values (consecutive n of one column m) go to LDS
// as one 8-byte store of the row-major activations.
// ---------------------------------------------------------------------------
constexpr int EM_NLEV = 20, EM_KIN = 64, EM_NH = 256, EM_NOUT = 21, EM_NOUTP = 32;
constexpr int EM_NIN = 2 * EM_NLEV + 3;  // features: T, ln q, T_s, ln p_s, ln CO2
constexpr int EM_MT = 32;             // columns per block
constexpr int EM_NT = 512;            // threads per block (8 waves)
constexpr int EM_LDH = EM_NH + 8;     // padded LDS row (bf16) of a hidden layer
constexpr int EM_LDX = EM_KIN + 8;

// this wave's weight fragments of one layer (output features [32w, 32w+32),
// K/32 x 2 x 16 bytes per lane), requested ahead of the layer that uses them
template <int K>
__device__ __forceinline__ void em_load(const uint16_t* W, V8 (&wa)[K / 32][2]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane >> 4, c = lane & 15;
  const uint16_t* wr = W + (long)(wv * 32 + c) * K + q * 8;
#pragma unroll
  for (int ks = 0; ks < K / 32; ++ks)
#pragma unroll
    for (int i = 0; i < 2; ++i) wa[ks][i].q = *(const uint4*)(wr + (long)i * 16 * K + ks * 32);
}

// one hidden layer: Hout[32][256] = gelu(Hin[32][K] W^T + b); wave w owns
// output features [32w, 32w + 32)
template <int K, int LDI>
__device__ __forceinline__ void em_hidden(const uint16_t* hin, const V8 (&wa)[K / 32][2],
                                          const float* b, uint16_t* hout) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane >> 4, c = lane & 15;
  const int n0 = wv * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < K / 32; ++ks) {
    V8 h[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
      h[j].q = *(const uint4*)(hin + (j * 16 + c) * LDI + ks * 32 + q * 8);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma(wa[ks][i], h[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int nb = n0 + i * 16 + 4 * q;  // this lane's 4 consecutive features
    const float4 bn = *(const float4*)(b + nb);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int m = j * 16 + c;
      uint2 pk;
      pk.x = (uint32_t)f2bf(gelu(acc[i][j][0] + bn.x)) |
             ((uint32_t)f2bf(gelu(acc[i][j][1] + bn.y)) << 16);
      pk.y = (uint32_t)f2bf(gelu(acc[i][j][2] + bn.z)) |
             ((uint32_t)f2bf(gelu(acc[i][j][3] + bn.w)) << 16);
      *(uint2*)(hout + m * EM_LDH + nb) = pk;
    }
  }
}

// GUARD: also flag (outside[col] = 1) every column with a feature outside
// [flo, fhi] (or not a number): the columns the network was not trained on
template <bool GUARD>
__global__ __launch_bounds__(EM_NT) void k_emulate(exo_lw_net net, int ncol, const float* dt,
                                                   const float* dq, const float* dp,
                                                   const float* co2, float* heat, float* olr,
                                                   const float* flo, const float* fhi,
                                                   int* outside) {
  __shared__ __align__(16) uint16_t sx[EM_MT * EM_LDX];
  __shared__ __align__(16) uint16_t sh0[EM_MT * EM_LDH];
  __shared__ __align__(16) uint16_t sh1[EM_MT * EM_LDH];
  __shared__ int sflag[EM_MT];
  const int tid = threadIdx.x;
  const int col0 = blockIdx.x * EM_MT;
  if (GUARD) {
    if (tid < EM_MT) sflag[tid] = 0;
    __syncthreads();
  }
  // every layer's weights are on their way while the features are computed
  // (one block per CU at a model step's 8192 columns: nothing else would
  // hide a later request's latency)
  const int lane = tid & 63, wv = tid >> 6;
  const int q = lane >> 4, c = lane & 15;
  const int nt = wv & 1, mt = (wv >> 1) & 1;
  V8 w1[EM_KIN / 32][2], w2[EM_NH / 32][2], w3[EM_NH / 32][2], w4[EM_NH / 32];
  em_load<EM_KIN>((const uint16_t*)net.w[0], w1);
  em_load<EM_NH>((const uint16_t*)net.w[1], w2);
  em_load<EM_NH>((const uint16_t*)net.w[2], w3);
  if (wv < 4) {
    const uint16_t* W4 = (const uint16_t*)net.w[3] + (long)(nt * 16 + c) * EM_NH + q * 8;
#pragma unroll
    for (int k = 0; k < EM_NH / 32; ++k) w4[k].q = *(const uint4*)(W4 + 32 * k);
  }
  // features (exoplasim_amd.radiation.features), column index fastest: all
  // of a thread's inputs are requested before any is used
  constexpr int NIT = EM_MT * EM_KIN / EM_NT;
  static_assert(EM_MT * EM_KIN % EM_NT == 0, "feature items per thread");
  float raw[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = tid + EM_NT * it;
    const int m = idx % EM_MT, f = idx / EM_MT;
    const int col = min(col0 + m, ncol - 1);
    const float* src = (f < EM_NLEV)           ? dt + (long)f * ncol
                       : (f < 2 * EM_NLEV)     ? dq + (long)(f - EM_NLEV) * ncol
                       : (f == 2 * EM_NLEV)    ? dt + (long)EM_NLEV * ncol
                       : (f == 2 * EM_NLEV + 1) ? dp
                                                : co2;
    raw[it] = (f < EM_NIN) ? src[col] : 0.f;
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = tid + EM_NT * it;
    const int m = idx % EM_MT, f = idx / EM_MT;
    const float x = raw[it];
    float v = 0.f;
    if (f < EM_NLEV)
      v = (x - 250.0f) / 50.0f;
    else if (f < 2 * EM_NLEV)
      v = (logf(fmaxf(x, 1e-12f)) + 10.0f) / 3.5f;
    else if (f == 2 * EM_NLEV)
      v = (x - 265.0f) / 40.0f;
    else if (f == 2 * EM_NLEV + 1)
      v = logf(x / 1e5f) / 1.3f;
    else if (f == 2 * EM_NLEV + 2)
      v = (logf(x) - 5.75f) / 3.3f;
    sx[m * EM_LDX + f] = f2bf(v);
    if (GUARD && f < EM_NIN && !(v >= flo[f] && v <= fhi[f])) atomicOr(&sflag[m], 1);
  }
  __syncthreads();
  if (GUARD && tid < EM_MT && col0 + tid < ncol) outside[col0 + tid] = sflag[tid];
  em_hidden<EM_KIN, EM_LDX>(sx, w1, net.b[0], sh0);
  __syncthreads();
  em_hidden<EM_NH, EM_LDH>(sh0, w2, net.b[1], sh1);
  __syncthreads();
  em_hidden<EM_NH, EM_LDH>(sh1, w3, net.b[2], sh0);
  __syncthreads();
  // output layer, transposed: waves 0..3 take feature tile wv&1 (n) and
  // column tile wv>>1 (m)
  if (wv >= 4) return;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < EM_NH / 32; ++k) {
    V8 h;
    h.q = *(const uint4*)(sh0 + (mt * 16 + c) * EM_LDH + 32 * k + q * 8);
    acc = mfma(w4[k], h, acc);
  }
  const int col = col0 + mt * 16 + c;
  if (col >= ncol) return;
  const float ps = dp[col] / 1e5f;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int n = nt * 16 + 4 * q + r;
    if (n >= EM_NOUT) continue;
    const float y = (acc[r] + net.b[3][n]) * net.ystd[n] + net.ymean[n];
    if (n < EM_NLEV)
      heat[(long)n * ncol + col] = y / ps;
    else
      olr[col] = y;
  }
}

}  // namespace

extern "C" int exo_mlp_forward(const void* X, const void* W, const float* b, int M, int K, int N,
                               int nvalid, int act, float* Z, void* H, float* Y, void* stream) {
  if (!X || !W || !b || M < 0 || K % 32 || N % 32 || (act && !H) || (!act && !Y))
    return EXO_RAD_ERR_ARG;
  if (M == 0) return EXO_RAD_OK;
  const int nt = N / 64 > 0 ? N / 64 : 1;
  if (N % 64 && N != 32) return EXO_RAD_ERR_ARG;
  k_fwd<<<nb(M, 64) * nt, 256, 0, (hipStream_t)stream>>>(
      (const uint16_t*)X, (const uint16_t*)W, b, M, K, N, nvalid, act, Z, (uint16_t*)H, Y);
  return hipGetLastError() == hipSuccess ? EXO_RAD_OK : EXO_RAD_ERR_HIP;
}

extern "C" int exo_mlp_backward_dx(const void* dZ, const void* Wt, const float* Zprev, int M,
                                   int N, int K, int kvalid, void* dZprev, float* dZprev32,
                                   void* stream) {
  if (!dZ || !Wt || !Zprev || !dZprev || M < 0 || K % 64 || N % 32) return EXO_RAD_ERR_ARG;
  if (M == 0) return EXO_RAD_OK;
  k_bwd_dx<<<nb(M, 64) * (K / 64), 256, 0, (hipStream_t)stream>>>(
      (const uint16_t*)dZ, (const uint16_t*)Wt, Zprev, M, N, K, kvalid, (uint16_t*)dZprev,
      dZprev32);
  return hipGetLastError() == hipSuccess ? EXO_RAD_OK : EXO_RAD_ERR_HIP;
}

extern "C" int exo_mlp_backward_dw(const void* dZ, const void* X, const float* dZ32, int M,
                                   int N, int K, int nsplit, float* part, float* dW, float* db,
                                   void* stream) {
  if (!dZ || !X || !dZ32 || !part || !dW || !db || M < 1 || nsplit < 1 || N % 32 || K % 32)
    return EXO_RAD_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int ntn = N / 64 > 0 ? N / 64 : 1, ntk = K / 64 > 0 ? K / 64 : 1;
  if ((N % 64 && N != 32) || (K % 64 && K != 32)) return EXO_RAD_ERR_ARG;
  int mchunk = (M + nsplit - 1) / nsplit;
  mchunk = ((mchunk + 31) / 32) * 32;
  const int ns = (M + mchunk - 1) / mchunk;
  k_bwd_dw<<<ntn * ntk * ns, 256, 0, s>>>((const uint16_t*)dZ, (const uint16_t*)X, M, N, K,
                                           mchunk, part);
  k_sum_dw<<<nb((long)N * K, 256), 256, 0, s>>>(part, ns, (long)N * K, dW);
  // the bias partials reuse `part` (ns x N <= ns x N x K floats) once the
  // weight partials are summed
  k_bias_grad<<<ntn * ns, 256, 0, s>>>(dZ32, M, N, mchunk, part);
  k_sum_dw<<<nb(N, 256), 256, 0, s>>>(part, ns, N, db);
  return hipGetLastError() == hipSuccess ? EXO_RAD_OK : EXO_RAD_ERR_HIP;
}

extern "C" int exo_mlp_cast_weights(const float* w, int nout, int nin, int N, int K, void* Wb,
                                    void* Wtb, void* stream) {
  if (!w || !Wb || !Wtb || nout > N || nin > K) return EXO_RAD_ERR_ARG;
  k_cast_w<<<nb((long)N * K, 256), 256, 0, (hipStream_t)stream>>>(w, nout, nin, N, K,
                                                                  (uint16_t*)Wb, (uint16_t*)Wtb);
  return hipGetLastError() == hipSuccess ? EXO_RAD_OK : EXO_RAD_ERR_HIP;
}

extern "C" int exo_lw_emulate(const exo_lw_net* net, int ncol, const float* dt, const float* dq,
                              const float* dp, const float* dqco2, float* heat, float* olr,
                              void* stream) {
  if (!net || ncol < 0 || net->nlev != EM_NLEV) return EXO_RAD_ERR_ARG;
  if (ncol == 0) return EXO_RAD_OK;
  if (!dt || !dq || !dp || !dqco2 || !heat || !olr || !net->ymean || !net->ystd)
    return EXO_RAD_ERR_ARG;
  for (int l = 0; l < 4; ++l)
    if (!net->w[l] || !net->b[l]) return EXO_RAD_ERR_ARG;
  k_emulate<false><<<nb(ncol, EM_MT), EM_NT, 0, (hipStream_t)stream>>>(
      *net, ncol, dt, dq, dp, dqco2, heat, olr, nullptr, nullptr, nullptr);
  return hipGetLastError() == hipSuccess ? EXO_RAD_OK : EXO_RAD_ERR_HIP;
}

extern "C" int exo_lw_emulate_guarded(const exo_lw_net* net, const float* flo, const float* fhi,
                                      int ncol, const float* dt, const float* dq,
                                      const float* dp, const float* dqco2, float* heat,
                                      float* olr, int* outside, void* stream) {
  if (!net || ncol < 0 || net->nlev != EM_NLEV) return EXO_RAD_ERR_ARG;
  if (ncol == 0) return EXO_RAD_OK;
  if (!dt || !dq || !dp || !dqco2 || !heat || !olr || !net->ymean || !net->ystd || !flo ||
      !fhi || !outside)
    return EXO_RAD_ERR_ARG;
  for (int l = 0; l < 4; ++l)
    if (!net->w[l] || !net->b[l]) return EXO_RAD_ERR_ARG;
  k_emulate<true><<<nb(ncol, EM_MT), EM_NT, 0, (hipStream_t)stream>>>(
      *net, ncol, dt, dq, dp, dqco2, heat, olr, flo, fhi, outside);
  return hipGetLastError() == hipSuccess ? EXO_RAD_OK : EXO_RAD_ERR_HIP;
}

extern "C" int exo_mlp_cast_rows(const float* src, int M, int nin, int K, void* dst,
                                 void* stream) {
  if (!src || !dst || nin > K || M < 0) return EXO_RAD_ERR_ARG;
  if (M == 0) return EXO_RAD_OK;
  k_cast_rows<<<nb((long)M * K, 256), 256, 0, (hipStream_t)stream>>>(src, M, nin, K,
                                                                     (uint16_t*)dst);
  return hipGetLastError() == hipSuccess ? EXO_RAD_OK : EXO_RAD_ERR_HIP;
}
