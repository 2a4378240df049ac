// sht_bodies.h -- the work items of the inverse and direct transforms as
// device functions, shared by the per-truncation kernels
// (sht_transforms.hip) and the persistent mixed-truncation ensemble kernels
// (sht_ensemble.hip).  Each body takes its block index and an LDS pointer,
// so one kernel can run items of different truncations in turn.
#pragma once
#include <type_traits>

#include "sht_internal.h"

namespace exo {

template <int MODE>
using RealT = typename std::conditional<MODE == MODE_F64, double, float>::type;

// ---------------------------------------------------------------------------
// tile shapes
// ---------------------------------------------------------------------------
// V selects the tile variant the autotuner times: V = 1 halves the fields
// per block (half the LDS, twice the blocks in flight); V = 2 keeps the
// fields and runs the block with 512 threads (8 waves: twice the waves in
// flight per CU for the same LDS, FFT batches of twice the rows; the staged
// T127 block 1024 threads); V = 3
// halves the fields and runs 512 threads (the persistent ensemble kernels,
// whose LDS is sized for their largest truncation).
template <int T, int MODE, int V = 0>
struct InvCfg {
  static constexpr int LP = 8;  // latitude pairs per tile (= 16 MFMA rows)
  // STAGED (T127, MFMA modes): the fused inverse keeps the Legendre sums as
  // compact Fourier coefficients (m <= T, 16 bytes per row and m: 128 KB for
  // 8 pairs x 8 fields, all 16 MFMA columns) and each wave then transforms
  // its rows one at a time in a row buffer of its own.  Full FFT rows of 8
  // fields (202 KB) would not fit a CU's 160 KB of LDS.  (The same layout at
  // T63, 78 KB and two blocks per CU, made the 32-member inverse 4% faster
  // but the round trip no faster and the ensemble 1.5% slower.)
  static constexpr bool STAGED = (T == 127) && (MODE != MODE_F64);
  // T63: 6 field-levels (77.6 KB of FFT rows on 512 threads, two blocks per
  // CU, 12 of the 16 MFMA columns) beat both 8 (103 KB, one block per CU)
  // and 4 (three blocks, half the columns)
  static constexpr int FB0 =
      (T == 127) ? (MODE == MODE_F64 ? 2 : 8) : (MODE == MODE_F64 ? 4 : (T == 63 ? 6 : 8));
  static constexpr int FB = (V == 1 || V == 3) ? ((FB0 >> 1) < 2 ? 2 : (FB0 >> 1)) : FB0;
  // threads per block; the staged V = 2 block runs 16 waves in the Legendre
  // phase, of which 8 then run the FFT rows (one row buffer each)
  static constexpr int NT = (V == 2 || V == 3) ? ((STAGED && V == 2) ? 1024 : 512) : 256;
  static constexpr int NWF = (NT / 64 < 8) ? NT / 64 : 8;  // staged: FFT waves
  // MFMA work items in flight per wave, one table (D1) or two (D2; dv2uv
  // carries two tables per item, and one item ahead keeps its registers
  // within four waves per SIMD)
  static constexpr int D1 = 2;
  static constexpr int D2 = 1;
  // Latitude-pair blocks per tile of the block order (scalar, dv2uv items):
  // a member's items run as PBT pair blocks x all its field groups, then the
  // next PBT pair blocks.  An XCD runs about 32 consecutive items at once,
  // which then read PBT table slices (4 x 442 KB at T127) instead of all
  // NPAIR / LP of them (5.3 MB per table at T127, above the XCD's 4 MB L2,
  // which a field-group-major order fetches again for every group).  Below
  // T127 all slices of a table fit in L2 together, and the order is field
  // group major (PBT = NPAIR / LP).
  static constexpr int PBT_SC = (T == 127) ? 4 : TS<T>::NPAIR / LP;
  static constexpr int PBT_UV = (T == 127) ? 6 : TS<T>::NPAIR / LP;
};

template <int T, int MODE, int V = 0>
struct FwdCfg {
  static constexpr int PB = (T == 21) ? 16
                          : (T == 42) ? 32
                          : (T == 63) ? (MODE == MODE_F64 ? 24 : 48)
                                      : (MODE == MODE_F64 ? 16 : 32);
  static constexpr int FBF0 = (T == 21) ? 4 : (T == 42) ? (MODE == MODE_F64 ? 1 : 2) : 1;
  static constexpr int FBF = (V == 1 || V == 3) ? ((FBF0 >> 1) < 1 ? 1 : (FBF0 >> 1)) : FBF0;
  static constexpr int NPB = TS<T>::NPAIR / PB;
  static constexpr int NT = (V == 2 || V == 3) ? 512 : 256;
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
// LDS of one inverse block.  Unfused: the Fourier coefficients of LP pairs
// x FB fields (north and south).  Fused: the FFT rows of those LP x FB
// (pair, field) rows, z_k = fc_N(k) + i fc_S(k) with its Hermitian extension,
// which the Legendre sums write directly, plus the twiddles.
// Staged (InvCfg::STAGED): the coefficients as unfused, then one padded FFT
// row per wave.
template <int T, int MODE, bool FUSED, int V>
struct InvLds {
  using R = RealT<MODE>;
  using C = InvCfg<T, MODE, V>;
  static constexpr bool ST = FUSED && C::STAGED;
  static constexpr size_t FCL =
      (FUSED && !ST) ? 0 : sizeof(NS2<R>) * C::LP * C::FB * TS<T>::NM;
  static constexpr size_t FBUF =
      ST      ? sizeof(C2<R>) * C::NWF * fft_buf_elems(TS<T>::NLON)
      : FUSED ? sizeof(C2<R>) * fft_buf_elems(C::LP * C::FB * TS<T>::NLON)
              : 0;
  static constexpr size_t TWS = FUSED ? sizeof(C2<R>) * TS<T>::NLON : 0;
  static constexpr size_t BYTES = FCL + FBUF + TWS;
};

// number of blocks (work items) of one inverse launch
template <int T, int MODE, int V>
__host__ __device__ inline long inv_items(const InvArgs& a) {
  using C = InvCfg<T, MODE, V>;
  const int ngroups = (a.kind == INV_UV) ? (a.nlev + C::FB / 2 - 1) / (C::FB / 2)
                                         : (a.nfield + C::FB - 1) / C::FB;
  return (long)a.nmem * ngroups * (TS<T>::NPAIR / C::LP);
}

// One work item of the inverse transform: pair block pb, field group g,
// member mem, decoded from blk.  lds holds InvLds::BYTES bytes.
template <int T, int MODE, int KIND, bool FUSED, int V>
__device__ __forceinline__ void inv_body(const InvArgs& a, int blk, const void* tabA0,
                                         const void* tabA1, long plane, const void* tabv,
                                         const void* twp, char* lds) {
  using S = TS<T>;
  using R = RealT<MODE>;
  constexpr int NM = S::NM, N = S::NLON, NLAT = S::NLAT, NPB = S::NPAIR / 8;
  constexpr int LP = InvCfg<T, MODE, V>::LP, FB = InvCfg<T, MODE, V>::FB;
  constexpr int NT = InvCfg<T, MODE, V>::NT, NW = NT / 64;  // threads, waves
  constexpr int NTERM = (KIND == INV_UV) ? 2 : 1;
  constexpr int NRP = N + N / 32;  // padded FFT row (fused)
  constexpr bool ST = FUSED && InvCfg<T, MODE, V>::STAGED;

  using L = InvLds<T, MODE, FUSED, V>;
  NS2<R>* fcl = (NS2<R>*)lds;
  C2<R>* fbuf = (C2<R>*)(lds + L::FCL);
  C2<R>* tws = (C2<R>*)(lds + L::FCL + L::FBUF);

  const int tid = item_tid();
  const int ngroups = (KIND == INV_UV) ? (a.nlev + FB / 2 - 1) / (FB / 2) : (a.nfield + FB - 1) / FB;
  // blk -> (member, pair block, field group): per member, tiles of PBT pair
  // blocks, each tile running its PBT pair blocks of field group 0, then of
  // group 1, ...
  constexpr int PBT =
      (KIND == INV_UV) ? InvCfg<T, MODE, V>::PBT_UV : InvCfg<T, MODE, V>::PBT_SC;
  static_assert(NPB % PBT == 0, "pair-block tiles must divide the pair blocks");
  const int mem = blk / (NPB * ngroups);
  if (mem >= a.nmem) return;
  const int loc = blk - mem * (NPB * ngroups);
  const int tile = loc / (PBT * ngroups), rt = loc - tile * (PBT * ngroups);
  const int g = rt / PBT;
  const int pb = tile * PBT + (rt - g * PBT);
  const R* spm = (const R*)a.sp + mem * a.sp_stride;
  R* outm = (R*)a.out + mem * a.out_stride;

  if (FUSED) {
    const C2<R>* tw = (const C2<R>*)twp;
    for (int k = tid; k < N; k += NT) tws[k] = tw[k];
    // FFT inputs T+1 .. N-T-1 (above the truncation, Nyquist included) are 0
    // (staged: not stored, the FFT's first pass leaves them out)
    constexpr int NZ = N - 2 * T - 1;
    if constexpr (!ST) {
      for (int i = tid; i < LP * FB * NZ; i += NT) {
        const int f = i / NZ, k = T + 1 + (i - f * NZ);
        fbuf[f * NRP + k + (k >> 5)] = C2<R>{R(0), R(0)};
      }
    }
  }

  // Fused: FFT row f = pl*FB + fl gets z_m = (n.x - s.y, n.y + s.x) and
  // z_{N-m} = (n.x + s.y, s.x - n.y) from the northern (n) and southern (s)
  // coefficients of wavenumber m; the imaginary parts of m = 0 are ignored,
  // as fc2gp does.  Unfused: the coefficients as they are.
  auto put = [&](int pl, int fl, int m, NS2<R> e) {
    if constexpr (FUSED && !ST) {
      C2<R>* row = fbuf + (pl * FB + fl) * NRP;
      if (m == 0) {
        row[0] = C2<R>{e.n.x, e.s.x};
      } else {
        const int kc = N - m;
        row[m + (m >> 5)] = C2<R>{e.n.x - e.s.y, e.n.y + e.s.x};
        row[kc + (kc >> 5)] = C2<R>{e.n.x + e.s.y, e.s.x - e.n.y};
      }
    } else {
      fcl[(pl * FB + fl) * NM + m] = e;
    }
  };

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
      if (fl >= FB || (fl >> 1) >= FB / 2 || lvl >= a.nlev) return R(0);
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

  // ---- phase 1: Legendre sums, into the FFT rows (fused) or fcl ----
  if constexpr (MODE == MODE_BF16X3 || MODE == MODE_F32MFMA || MODE == MODE_F64) {
    // GEMM per m: C[row = 2*pair + class][col] = sum_n A[row][n] B[n][col];
    // split-BF16 on v_mfma_f32_16x16x32_bf16, FP32 on v_mfma_f32_16x16x4_f32,
    // FP64 on v_mfma_f64_16x16x4_f64 (the FP32 operand layouts in doubles);
    // wave wv of NW takes m = wv, wv + NW, ...; each (m, 32-wide n block) is
    // one work item, and the operands of the items ahead are fetched before
    // the current one is multiplied.
    // The wave index is read into a scalar register, so that the item loop
    // (m, kb), its branches and the operand offsets are wave-uniform scalar
    // work rather than per-lane vector arithmetic.
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int q = lane >> 4, c = lane & 15;
    // column c of B: SCALAR field g*FB + c/2; UV level g*FB/2 + c/4 of
    // vorticity ((c/2)&1 == 0) or divergence; component c&1
    const R* bp = nullptr;  // F32MFMA, F64: the column's coefficients
    bool isvor = false;
    int pf = -1;  // field of column c
    if (KIND == INV_SCALAR) {
      const int field = g * FB + (c >> 1);
      if ((c >> 1) < FB && field < a.nfield) pf = a.src0 + field;
    } else {
      const int lvl = g * (FB / 2) + (c >> 2);
      const int s = (c >> 1) & 1;
      if ((c >> 2) < FB / 2 && lvl < a.nlev) {
        pf = (s ? a.div0 : a.vor0) + lvl;
        isvor = (s == 0) && ((c & 1) == 0);
      }
    }
    if constexpr (MODE != MODE_BF16X3) {
      if (pf >= 0) bp = spm + (long)pf * S::NRSP + (c & 1);
    }
    // Operands through buffer loads: a per-lane byte offset plus a
    // wave-uniform byte offset per item.
    // BF16X3: the K index of lane quarter q is parity h = q >> 1, row
    // element j = 16 kb + 8 (q & 1) .. + 7 (layouts in sht_common.h).
    // A: the table rows of pair block pb; a lane whose parity is not its
    // row's class, or whose j lies past the row, reads past the buffer,
    // which returns the zeros.  F32MFMA, F64: item (m, kb) at ((m KB + kb)
    // 512) elements of the block's table.
    constexpr int AE = (MODE == MODE_BF16X3) ? 2 : (int)sizeof(R);  // bytes per table element
    const long abase = (MODE == MODE_BF16X3) ? (long)pb * tab_pb(T) : (long)pb * NM * S::KB * 512;
    const int abytes = (MODE == MODE_BF16X3) ? tab_pb(T) * 2 : (int)(plane - abase) * AE;
    const __amdgpu_buffer_rsrc_t rsA0 = buffer_rsrc((const char*)tabA0 + abase * AE, abytes);
    const __amdgpu_buffer_rsrc_t rsA1 =
        buffer_rsrc((const char*)(NTERM == 2 ? tabA1 : tabA0) + abase * AE, abytes);
    const int hq = q >> 1, jq = 8 * (q & 1);
    // does this lane hold nonzero table values (its parity is its row's class)
    const bool amine0 = hq == ((c & 1) ^ (table_is_q(KIND == INV_UV ? TQV : a.table) ? 1 : 0));
    const bool amine1 = hq == ((c & 1) ^ (table_is_q(TQU) ? 1 : 0));
    const int avoff = (q * 16 + c) * 8 * AE;  // F32MFMA, F64: 8 values per lane
    // B (BF16X3): the member's parity-ordered fp32 operand; columns without
    // a field read past the buffer, which returns zeros
    __amdgpu_buffer_rsrc_t rsB = rsA0;
    int bvoff = BUF_NONE;
    if constexpr (MODE == MODE_BF16X3) {
      rsB = buffer_rsrc((const float*)a.bpk + mem * a.bpk_stride, (int)a.bpk_stride * 4);
      if (pf >= 0) bvoff = ((pf - a.bpk_f0) * pk_field(T) + jq) * 4;
    }
    const int brow = (c & 1) * 2 + hq;  // re/im, parity row of the B segment
    const R pv = (R)a.plavor;
    // B fragment of one item: 8 values (F32MFMA, F64: f, n = 4j + q), or 8
    // fp32 values of one parity as loaded (BF16X3: r, split into the three
    // bf16 planes when the item is multiplied; fix marks the lane that holds
    // the absolute vorticity at (m=0, n=1, re))
    struct BFrag {
      R f[8];
      uint4 r[2];
      bool fix;
    };
    // BF16X3 table planes over lane pairs (two loads per item and term, in
    // fetch below) from T63 up: there the load instructions bound the
    // Legendre phase (T63 inverse 5% faster); at T21 and T42 the pair
    // exchange's VALU work cost more than the load it saves (T42 inverse
    // 1-3% slower)
    constexpr bool PAIRLOAD = (MODE == MODE_BF16X3) && T >= 63;
    // 16-byte table loads per item and term: three bf16 planes (two with
    // PAIRLOAD), or 8 values of 4 or 8 bytes
    constexpr int NA = (MODE == MODE_BF16X3) ? (PAIRLOAD ? 2 : 3) : (8 * AE) / 16;
    // Every step fetches, also past the last item (m > T: nothing is loaded,
    // at the offsets of m = T), so that each step issues the same loads and
    // the wait for an item's operands counts only the loads issued before
    // them.
    auto fetch = [&](int mi, int kb, BFrag& br, uint4* a0, uint4* a1) {
      if constexpr (MODE == MODE_BF16X3) {
        const bool live = mi <= T;
        const int m = live ? mi : T;
        const int rw = pk_rw(T, m);
        // lanes whose j lies past the row would read n > NTRU, where the
        // table operand is zero: they load nothing
        const bool inrow = live && 16 * kb + jq < rw;
        const int bso = (pk_seg(T, m) + 16 * kb) * 4;
        const int bv = inrow ? bvoff + brow * rw * 4 : BUF_NONE;
        br.r[0] = buffer_load_x4(rsB, bv, bso);
        br.r[1] = buffer_load_x4(rsB, bv, bso + 16);
        // absolute -> relative vorticity at (m=0, n=1, re), the first odd
        // value (subtracted when the item is multiplied)
        br.fix = live && m == 0 && kb == 0 && isvor && q == 2;
        // PAIRLOAD: rows c and c ^ 1 are the two classes of one latitude
        // pair, so of the lanes c and c ^ 1 of a quarter exactly one holds
        // nonzero values.  That lane loads its hi plane (first load) and lo
        // plane (second); the other lane loads the partner's mid plane in
        // the first load, which the pair exchanges when the item is
        // multiplied: two 16-byte loads per term instead of three (a wave's
        // load costs the same however many of its lanes read).  Else each
        // lane loads its three planes, the other class's lanes nothing.
        const int aso = (tab_seg(T, m) + 16 * kb) * 2;
        const int av = (c * rw + jq) * 2;                          // own row, hi plane
        if constexpr (!PAIRLOAD) {
          const int av0 = (amine0 && inrow) ? av : BUF_NONE;
#pragma unroll
          for (int p = 0; p < 3; ++p) a0[p] = buffer_load_x4(rsA0, av0, aso + p * rw * 32);
          if (NTERM == 2) {
            const int av1 = (amine1 && inrow) ? av : BUF_NONE;
#pragma unroll
            for (int p = 0; p < 3; ++p) a1[p] = buffer_load_x4(rsA1, av1, aso + p * rw * 32);
          }
          return;
        }
        const int avp = ((c ^ 1) * rw + jq) * 2 + rw * 32;         // partner's row, mid plane
        const int avl = av + rw * 64;                              // own row, lo plane
        {
          const int v1 = inrow ? (amine0 ? av : avp) : BUF_NONE;
          const int v2 = (amine0 && inrow) ? avl : BUF_NONE;
          a0[0] = buffer_load_x4(rsA0, v1, aso);
          a0[1] = buffer_load_x4(rsA0, v2, aso);
        }
        if (NTERM == 2) {
          const int v1 = inrow ? (amine1 ? av : avp) : BUF_NONE;
          const int v2 = (amine1 && inrow) ? avl : BUF_NONE;
          a1[0] = buffer_load_x4(rsA1, v1, aso);
          a1[1] = buffer_load_x4(rsA1, v2, aso);
        }
      } else {
        // m > T: n > T loads nothing, and the table offset lies past the
        // table's buffer
        const int m = mi;
        const int w0 = w0_of(T, m <= T ? m : T);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int n = m + kb * 32 + j * 4 + q;
          R x = R(0);
          if (bp && n <= T) {
            x = bp[2 * (w0 + n - m)];
            if (isvor && m == 0 && n == 1) x = x - pv;
          }
          br.f[j] = x;
        }
        // operand stored as [q][row][8] so one lane reads its 8 values
        const int aso = (m * S::KB + kb) * 512 * AE;
#pragma unroll
        for (int i = 0; i < NA; ++i) a0[i] = buffer_load_x4(rsA0, avoff, aso + 16 * i);
        if (NTERM == 2) {
#pragma unroll
          for (int i = 0; i < NA; ++i) a1[i] = buffer_load_x4(rsA1, avoff, aso + 16 * i);
        }
      }
    };
    // BF16X3: the exact three-way bf16 split of an item's 8 B values
    auto split = [&](const BFrag& br, BF8* bp) {
      if constexpr (MODE == MODE_BF16X3) {
        float x[8] = {__uint_as_float(br.r[0].x), __uint_as_float(br.r[0].y),
                      __uint_as_float(br.r[0].z), __uint_as_float(br.r[0].w),
                      __uint_as_float(br.r[1].x), __uint_as_float(br.r[1].y),
                      __uint_as_float(br.r[1].z), __uint_as_float(br.r[1].w)};
        if (br.fix) x[0] = x[0] - (float)pv;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) split3(x[jj], bp[0].s[jj], bp[1].s[jj], bp[2].s[jj]);
      }
    };
    // two independent accumulation chains per term (the small cross terms
    // and the large ones), so that consecutive MFMAs do not all wait for the
    // previous result
    using Acc = typename std::conditional<MODE == MODE_F64, f64x4, f32x4>::type;
    auto mul = [&](const BFrag& br, const BF8* bp, const uint4* a, bool amine, Acc& acc,
                   Acc& acc2) {
      if constexpr (MODE == MODE_BF16X3) {
        // the lane holding its row's values: hi and lo as loaded, mid from
        // the partner lane (lane ^ 1); the other lane: zeros
        BF8 ah, am, al;
        if constexpr (PAIRLOAD) {
          const uint4 z = {0u, 0u, 0u, 0u};
          const uint4 pm = {quad_perm_u<0xB1>(a[0].x), quad_perm_u<0xB1>(a[0].y),
                            quad_perm_u<0xB1>(a[0].z), quad_perm_u<0xB1>(a[0].w)};
          ah.q = amine ? a[0] : z;
          am.q = amine ? pm : z;
          al.q = a[1];
        } else {
          ah.q = a[0];
          am.q = a[1];
          al.q = a[NA - 1];
        }
        acc = mfma_bf16(ah, bp[2], acc);
        acc2 = mfma_bf16(am, bp[0], acc2);
        acc = mfma_bf16(am, bp[1], acc);
        acc2 = mfma_bf16(ah, bp[1], acc2);
        acc = mfma_bf16(al, bp[0], acc);
        acc2 = mfma_bf16(ah, bp[0], acc2);
      } else if constexpr (MODE == MODE_F64) {
        const double* ad = (const double*)a;
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ad[j], br.f[j], acc, 0, 0, 0);
          acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ad[j + 1], br.f[j + 1], acc2, 0, 0, 0);
        }
      } else {
        const float* af = (const float*)a;
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[j], br.f[j], acc, 0, 0, 0);
          acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(af[j + 1], br.f[j + 1], acc2, 0, 0, 0);
        }
      }
    };
    // operands of one work item (m, kb)
    struct Frag {
      BFrag b;
      uint4 a0[NA], a1[NA];
    };
    struct Item {
      int m, kb;
    };
    // (read back as wave-uniform scalars: a select that the compiler makes
    // per lane would turn the table offsets into vector values and each load
    // into a loop over lanes)
    auto next_item = [&](Item it) -> Item {
      const int kbn = (NM - it.m + 31) >> 5;  // <= 0 past the truncation
      const Item n = (it.kb + 1 < kbn) ? Item{it.m, it.kb + 1} : Item{it.m + NW, 0};
      return Item{__builtin_amdgcn_readfirstlane(n.m), __builtin_amdgcn_readfirstlane(n.kb)};
    };
    const Acc zero4 = {0, 0, 0, 0};
    Acc acc1 = zero4, acc1b = zero4, acc2 = zero4, acc2b = zero4;
    // multiply item `cur` (operands in `use`) after starting the loads of
    // item `ahead` (two items ahead) into `into`; at the last n block of an
    // m, combine the terms, fold N/S and store the Fourier coefficients
    auto step = [&](Item cur, Frag& use, Frag& into, Item ahead) {
      fetch(ahead.m, ahead.kb, into.b, into.a0, into.a1);
      BF8 bp[3];
      split(use.b, bp);
      mul(use.b, bp, use.a0, amine0, acc1, acc1b);
      if (NTERM == 2) mul(use.b, bp, use.a1, amine1, acc2, acc2b);
      // (a step past the wave's last item multiplies the zeros it loaded
      // and stores nothing)
      if (next_item(cur).kb == 0 && cur.m <= T) {
        const int m = cur.m;
        acc1 += acc1b;
        if (NTERM == 2) acc2 += acc2b;
        Acc out = acc1;
        if (KIND == INV_UV) {
          // u = qv.zeta + qu.(-i D), v = qu.(-i zeta) - qv.D: the qu product of
          // the partner column (zeta.re <-> D.im, zeta.im <-> D.re)
          Acc p2;
#pragma unroll
          for (int i = 0; i < 4; ++i) p2[i] = quad_perm<0x1B>(acc2[i]);  // lane ^ 3
          // column c & 3 = 0: acc1 + p2, 1: acc1 - p2, 2: p2 - acc1, 3: -p2 - acc1
          const R sa = (c & 2) ? R(-1) : R(1), sb = (c & 1) ? R(-1) : R(1);
#pragma unroll
          for (int i = 0; i < 4; ++i) out[i] = sa * acc1[i] + sb * p2[i];
        }
        const int fl = c >> 1, ri = c & 1;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int pl = 2 * q + h;
          const R sy = out[2 * h], an = out[2 * h + 1];
          // component ri of the northern and southern coefficient
          R nv = sy + an, sv = sy - an;
          if (KIND == INV_SCALAR && a.dlam) {
            // d/dlambda: (x + i y) i m = -m y + i m x, the other component
            // from the partner column (lane ^ 1)
            const R fm = (R)m;
            const R np = quad_perm<0xB1>(nv), spn = quad_perm<0xB1>(sv);
            nv = ri ? np * fm : -np * fm;
            sv = ri ? spn * fm : -spn * fm;
          }
          if constexpr (FUSED && !ST) {
            // the partner column (lane ^ 1) holds the other component: sp
            // is s.y for ri = 0 and s.x for ri = 1
            const R sp = quad_perm<0xB1>(sv);  // lane ^ 1
            if (fl < FB) {
              R* row = (R*)(fbuf + (pl * FB + fl) * NRP);
              if (m == 0) {
                row[ri] = ri ? sp : nv;  // z_0 = (n.x, s.x)
              } else {
                const int kc = N - m;
                // z_m = (n.x - s.y, n.y + s.x), z_{N-m} = (n.x + s.y, s.x - n.y)
                row[2 * (m + (m >> 5)) + ri] = ri ? nv + sp : nv - sp;
                row[2 * (kc + (kc >> 5)) + ri] = ri ? sp - nv : nv + sp;
              }
            }
          } else {
            if (fl < FB) {
              R* e = (R*)&fcl[(pl * FB + fl) * NM + m];
              e[ri] = nv;
              e[2 + ri] = sv;
            }
          }
        }
        acc1 = acc1b = acc2 = acc2b = zero4;
      }
    };
    // DEPTH + 1 operand buffers in rotation, unrolled so that no buffer is
    // copied: DEPTH items are in flight while one is multiplied.  The step
    // of round r multiplies buffer r and fetches DEPTH items ahead into
    // buffer (r + DEPTH) mod (DEPTH + 1).
    constexpr int DEPTH = (NTERM == 1) ? InvCfg<T, MODE, V>::D1 : InvCfg<T, MODE, V>::D2;
    Frag fr[DEPTH + 1];
    Item cur = {wv, 0}, ahead = cur;
    // The items ahead load in item order, as the loop's steps load them:
    // the waits in the loop are counted once for the loop entry and the
    // back edge together, so a prologue whose last loads were the first
    // item's would make every pass of the loop wait for all its loads.
#pragma unroll
    for (int k = 0; k < DEPTH; ++k) {
      fetch(ahead.m, ahead.kb, fr[k].b, fr[k].a0, fr[k].a1);
      __builtin_amdgcn_sched_barrier(0);
      ahead = next_item(ahead);
    }
    // Whole rounds of DEPTH + 1 steps, the steps past the wave's last item
    // only issuing their (empty) loads: the loop has no exit between its
    // rounds, so its head is reached only from its start and its back edge,
    // with the same loads pending from both, and the waits in it count only
    // the older items.
    int nit = 0;
    for (int m = wv; m <= T; m += NW) nit += (NM - m + 31) >> 5;
    const int nround = __builtin_amdgcn_readfirstlane((nit + DEPTH) / (DEPTH + 1));
    for (int rd = 0; rd < nround; ++rd) {
#pragma unroll
      for (int r = 0; r <= DEPTH; ++r) {
        step(cur, fr[r], fr[(r + DEPTH) % (DEPTH + 1)], ahead);
        cur = next_item(cur);
        ahead = next_item(ahead);
      }
    }
  } else {
    // vector-ALU Legendre: one (pair, field, m) per thread, both components
    const R* tv = (const R*)tabv;
    const long tsz = (long)S::NPAIR * S::NCSP;
    constexpr int NPC = LP * FB;
    const int t0 = (KIND == INV_UV) ? TQV : a.table;
    const int t1 = TQU;
    for (int idx = tid; idx < NPC * NM; idx += NT) {
      const int pc = idx % NPC, m = idx / NPC;
      const int pl = pc % LP, fl = pc / LP;
      const int pair = pb * LP + pl;
      const int w0 = w0_of(T, m);
      R sy[2] = {0, 0}, an[2] = {0, 0};
      for (int n = m; n <= T; ++n) {
        const int w = w0 + n - m;
        const bool even = ((n - m) & 1) == 0;
#pragma unroll
        for (int term = 0; term < NTERM; ++term) {
          const int t = term == 0 ? t0 : t1;
          const R p = tv[t * tsz + (long)pair * S::NCSP + w];
#pragma unroll
          for (int ri = 0; ri < 2; ++ri) {
            const R x = bval(term, 2 * fl + ri, w);
            if (even != table_is_q(t))
              sy[ri] += p * x;
            else
              an[ri] += p * x;
          }
        }
      }
      NS2<R> e{{sy[0] + an[0], sy[1] + an[1]}, {sy[0] - an[0], sy[1] - an[1]}};
      if (KIND == INV_SCALAR && a.dlam) {
        // d/dlambda: (x + i y) i m = -m y + i m x
        const R fm = (R)m;
        e = NS2<R>{{-e.n.y * fm, e.n.x * fm}, {-e.s.y * fm, e.s.x * fm}};
      }
      put(pl, fl, m, e);
    }
  }
  __syncthreads();

  // destination field of FFT-field fl (-1 if not present)
  auto dst_of = [&](int fl) -> int {
    if (KIND == INV_SCALAR) {
      const int field = g * FB + fl;
      return field < a.nfield ? a.dst0 + field : -1;
    }
    // a block holds FB / 2 levels (u and v rows); with FB odd its last row
    // holds no field and must not reach the next group's first level
    if ((fl >> 1) >= FB / 2) return -1;
    const int lvl = g * (FB / 2) + (fl >> 1);
    if (lvl >= a.nlev) return -1;
    return ((fl & 1) ? a.v0 : a.u0) + lvl;
  };

  if constexpr (!FUSED) {
    // Fourier output fc(2, NLON/2, NLAT), zero above the truncation
    for (int idx = tid; idx < LP * FB * (N / 2); idx += NT) {
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
  } else if constexpr (ST) {
    // ---- phase 2 (staged): wave wv < NWF transforms rows wv, wv + NWF, ...
    // one at a time in its own row buffer (the other waves of a 16-wave
    // block have no row buffer and wait at the block barrier); the first FFT
    // pass builds z = fc_N + i fc_S with its Hermitian extension from the
    // compact coefficients as it reads them.  Re -> north row, Im -> south
    // row ----
    static_assert(N == 192 || N == 384, "the staged inverse runs the wave FFT of 192 or 384");
    constexpr int NWF = InvCfg<T, MODE, V>::NWF;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    C2<R>* wrow = fbuf + (wv < NWF ? wv : 0) * NRP;
    typedef R R4 __attribute__((ext_vector_type(4)));
    for (int f = wv; wv < NWF && f < LP * FB; f += NWF) {
      const int pl = f / FB, fl = f - pl * FB;
      const int d = dst_of(fl);
      if (d < 0) continue;
      const NS2<R>* e = fcl + f * NM;
      // input k = lane + 64 r: k <= T the stored m = k (z_0 = (n.x, s.x),
      // the imaginary parts of m = 0 ignored as fc2gp does); T < k < N - T
      // zero (above the truncation; whole passes r of it known at compile
      // time); else z_{N-m}, m = N - k
      auto load = [&](int r, int k) -> C2<R> {
        if (64 * r + 63 <= T || (64 * r <= T && k <= T)) {
          const NS2<R> v = e[k];
          return (k == 0) ? C2<R>{v.n.x, v.s.x} : C2<R>{v.n.x - v.s.y, v.n.y + v.s.x};
        }
        if (64 * r + 63 < N - T || k < N - T) return C2<R>{R(0), R(0)};
        const NS2<R> v = e[N - k];
        return C2<R>{v.n.x + v.s.y, v.s.x - v.n.y};
      };
      fft_wave<N, +1, R>(wrow, tws, lane, load);
      // whole rows in runs of 4 values per lane, non-temporal (as below)
      const int lat = pb * LP + pl;
      R* o = outm + (long)d * NLAT * N;
      for (int k = 4 * lane; k < N; k += 256) {
        const C2<R>* z = wrow + k + (k >> 5);
        const C2<R> z0 = z[0], z1 = z[1], z2 = z[2], z3 = z[3];
        __builtin_nontemporal_store(R4{z0.x, z1.x, z2.x, z3.x}, (R4*)(o + (long)lat * N + k));
        __builtin_nontemporal_store(R4{z0.y, z1.y, z2.y, z3.y},
                                    (R4*)(o + (long)(NLAT - 1 - lat) * N + k));
      }
      // the row is read before the wave's next row rewrites it
      wave_lds_sync();
    }
    // the coefficients are read before a next work item of the block
    // rewrites them
    __syncthreads();
  } else {
    // ---- phase 2: inverse FFT of the rows z = fc_N + i fc_S, all LP*FB at
    // once and in place: Re -> north row, Im -> south row ----
    auto none = [](int, int, C2<R>) {};
    fft_run_lds<N, +1, R, NT, LP * FB, T + 1, N - T - 1, false>(fbuf, tws, none);
    // the grid rows leave in runs of 4 values per lane, so that a wave
    // writes whole rows: non-temporal, so that they do not evict the
    // spectral operands and tables that the other blocks reuse from L2
    typedef R R4 __attribute__((ext_vector_type(4)));
    constexpr int NK4 = N / 4;
    for (int i = tid; i < LP * FB * NK4; i += NT) {
      const int f = i / NK4, k = (i - f * NK4) * 4;
      const int pl = f / FB, fl = f % FB;
      const int d = dst_of(fl);
      if (d < 0) continue;
      // k .. k+3 lie in one 32-element run of the padded row
      const C2<R>* z = fbuf + f * NRP + k + (k >> 5);
      const C2<R> z0 = z[0], z1 = z[1], z2 = z[2], z3 = z[3];
      const int lat = pb * LP + pl;
      R* o = outm + (long)d * NLAT * N + k;
      __builtin_nontemporal_store(R4{z0.x, z1.x, z2.x, z3.x}, (R4*)(o + (long)lat * N));
      __builtin_nontemporal_store(R4{z0.y, z1.y, z2.y, z3.y},
                                  (R4*)(o + (long)(NLAT - 1 - lat) * N));
    }
    // the rows are read before a next work item of the block rewrites them
    __syncthreads();
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
struct FwdLds {
  using R = RealT<MODE>;
  using C = FwdCfg<T, MODE, V>;
  // one (m, class) row of F: PB pairs x NCOL columns plus one pad of NCOL
  // values, so that the rows that one wave reads start on different banks
  static constexpr int RS = (C::PB + 1) * 2 * C::FBF;
  static constexpr size_t F = sizeof(R) * TS<T>::NM * 2 * RS;
  static constexpr size_t FBUF =
      FUSED ? sizeof(C2<R>) * fft_buf_elems(FftBatch<TS<T>::NLON, C::NT>::NF * TS<T>::NLON) : 0;
  static constexpr size_t TWS = FUSED ? sizeof(C2<R>) * TS<T>::NLON : 0;
  static constexpr size_t BYTES = F + FBUF + TWS;
};

template <int T, int MODE, int V>
__host__ __device__ inline long fwd_items(const FwdArgs& a) {
  using C = FwdCfg<T, MODE, V>;
  return (long)a.nmem * ((a.nfield + C::FBF - 1) / C::FBF) * C::NPB;
}

template <int T, int MODE, bool FUSED, int V>
__device__ __forceinline__ void fwd_body(const FwdArgs& a, int blk, const void* tabv,
                                         const void* twp, void* scratch, char* lds) {
  using S = TS<T>;
  using R = RealT<MODE>;
  using C = FwdCfg<T, MODE, V>;
  constexpr int NM = S::NM, N = S::NLON, NLAT = S::NLAT;
  constexpr int PB = C::PB, FBF = C::FBF, NPB = C::NPB;
  constexpr int NCOL = 2 * FBF;
  constexpr int NT = C::NT;
  constexpr int NF = FftBatch<N, NT>::NF;

  // F[m][cls][pair][col]: cls 0 = fc_N + fc_S (E), cls 1 = fc_N - fc_S (O)
  using L = FwdLds<T, MODE, FUSED, V>;
  R* F = (R*)lds;
  C2<R>* fbuf = (C2<R>*)(lds + L::F);
  C2<R>* tws = (C2<R>*)(lds + L::F + L::FBUF);

  const int tid = item_tid();
  const int ngroups = (a.nfield + FBF - 1) / FBF;
  const int pb = blk % NPB;
  const int g = (blk / NPB) % ngroups;
  const int mem0 = blk / (NPB * ngroups);
  if (mem0 >= a.nmem) return;
  // members in reverse order: in a round trip the inverse has just written
  // the grids, and its last members are the likeliest to be still cached
  const int mem = a.nmem - 1 - mem0;
  const R* inm = (const R*)a.in + mem * a.in_stride;

  // Rows pl x field of the tile: row gi holds pair pl = gi / FBF, field
  // g FBF + gi % FBF; z_k = g_N(k) + i g_S(k) of the northern and southern
  // latitudes of that pair (zero for fields past nfield).
  auto grid_z = [&](int gi, int k) -> C2<R> {
    const int pl = gi / FBF, fl = gi % FBF;
    const int field = g * FBF + fl;
    if (field >= a.nfield) return C2<R>{R(0), R(0)};
    const int lat = pb * PB + pl;
    const R* src = inm + (long)(a.src0 + field) * NLAT * N;
    return C2<R>{src[(long)lat * N + k], src[(long)(NLAT - 1 - lat) * N + k]};
  };

  // When the first FFT pass of every batch fits one butterfly per thread,
  // each thread requests its butterfly's inputs here, so that the grid
  // rows of all batches are in flight together
  constexpr int R0 = FftPlan<N>::R[0], NB0 = N / R0;
  constexpr int NBATCH = (PB * FBF + NF - 1) / NF;
  constexpr bool HELD = FUSED && NBATCH * NF * NB0 <= NT;
  C2<R> held[R0];
  const int hb = tid % (NF * NB0);               // butterfly within its batch
  const int hbatch = tid / (NF * NB0);           // batch of that butterfly
  const int hrow = hbatch * NF + hb / NB0;       // tile row
  if constexpr (HELD) {
    if (hrow < PB * FBF) {
      const int j = hb % NB0;
#pragma unroll
      for (int r = 0; r < R0; ++r) held[r] = grid_z(hrow, j + r * NB0);
    }
  }

  if (FUSED) {
    const C2<R>* tw = (const C2<R>*)twp;
    for (int k = tid; k < N; k += NT) tws[k] = C2<R>{tw[k].x, -tw[k].y};
  }
  __syncthreads();

  auto fidx = [](int m, int cls, int pl, int col) { return (m * 2 + cls) * L::RS + pl * NCOL + col; };

  if constexpr (HELD) {
    for (int bt = 0; bt < NBATCH; ++bt) {
      const int f0 = bt * NF;
      const int nf = min(NF, PB * FBF - f0);
      auto store = [&](int f, int k, C2<R> v) { fbuf[bix(f * N + k)] = v; };
      fft_pass0_held<N, -1, R>(fbuf, hbatch == bt && hrow < PB * FBF, hb, held);
      fft_run_after0<N, -1, R, NT, NF>(fbuf, tws, store, nf);
      // separate the two real rows, scale by 1/N, fold into E/O
      const R sc = R(0.5) / R(N);
      for (int idx = tid; idx < nf * NM; idx += NT) {
        const int f = idx % nf, k = idx / nf;
        const C2<R> zk = fbuf[bix(f * N + k)];
        const C2<R> zn = fbuf[bix(f * N + ((N - k) % N))];
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
  } else if constexpr (FUSED) {
    for (int f0 = 0; f0 < PB * FBF; f0 += NF) {
      const int nf = min(NF, PB * FBF - f0);
      auto load = [&](int f, int k) -> C2<R> { return grid_z(f0 + f, k); };
      auto store = [&](int f, int k, C2<R> v) { fbuf[bix(f * N + k)] = v; };
      fft_run<N, -1, R, NT>(fbuf, tws, nf, load, store);
      // separate the two real rows, scale by 1/N, fold into E/O
      const R sc = R(0.5) / R(N);
      for (int idx = tid; idx < nf * NM; idx += NT) {
        const int f = idx % nf, k = idx / nf;
        const C2<R> zk = fbuf[bix(f * N + k)];
        const C2<R> zn = fbuf[bix(f * N + ((N - k) % N))];
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
    for (int idx = tid; idx < PB * FBF * NM; idx += NT) {
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
  // FP64: the table [t][pair][w], one 8-byte load per pair; FP32: [t][pair /
  // 4][w][4], one 16-byte load per 4 pairs (a wave's loads cost the same
  // whatever their width, so the wider load reads 4x the bytes per load)
  static_assert(sizeof(R) == 8 || PB % 4 == 0, "pair blocks of whole groups of 4");
  const R* tv = (const R*)tabv + (long)a.table * S::NPAIR * S::NCSP + (long)pb * PB * S::NCSP;
  const __amdgpu_buffer_rsrc_t trs = buffer_rsrc(tv, PB * S::NCSP * (int)sizeof(R));
  for (int w = tid; w < S::NCSP; w += NT) {
    const int m = m_of_w(T, w);
    const int n = m + (w - w0_of(T, m));
    const int cls = ((((n - m) & 1) == 0) != qtab) ? 0 : 1;
    const R* fr = &F[fidx(m, cls, 0, 0)];
    R acc[NCOL];
#pragma unroll
    for (int c = 0; c < NCOL; ++c) acc[c] = R(0);
    if constexpr (sizeof(R) == 8) {
#pragma unroll 8
      for (int pl = 0; pl < PB; ++pl) {
        const R p = buffer_load<R>(trs, w * (int)sizeof(R), pl * S::NCSP * (int)sizeof(R));
#pragma unroll
        for (int c = 0; c < NCOL; ++c) acc[c] += p * fr[pl * NCOL + c];
      }
    } else {
#pragma unroll 2
      for (int pq = 0; pq < PB / 4; ++pq) {
        const uint4 p4 = buffer_load_x4(trs, w * 16, pq * S::NCSP * 16);
        const R p[4] = {__uint_as_float(p4.x), __uint_as_float(p4.y), __uint_as_float(p4.z),
                        __uint_as_float(p4.w)};
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int c = 0; c < NCOL; ++c) acc[c] += p[u] * fr[(4 * pq + u) * NCOL + c];
      }
    }
#pragma unroll
    for (int c = 0; c < NCOL; ++c) {
      const int field = g * FBF + (c >> 1);
      if (field < a.nfield) outp[(long)field * S::NRSP + 2 * w + (c & 1)] = acc[c];
    }
  }
}

}  // namespace exo
