// sht_internal.h -- library-internal declarations (plan object, kernels,
// launch helpers).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "sht_common.h"

struct exo_sht_plan_s;

namespace exo {

__global__ void k_inigau(int klat, double* pz0, double* pzw);
__global__ void k_legini(int ntru, int npair, const double* sid, const double* gwd, double* tab);
__global__ void k_filtered(int ntru, int npair, const double* tab, const double* skgpsp,
                           const double* skspgp, double* out64, float* out32);
__global__ void k_pack_inv(int ntru, int npair, int kbmax, int t, const double* filt64,
                           float* f32, double* f64);
__global__ void k_pack_inv_bf(int ntru, int npair, int t, const double* filt64, uint16_t* bf);
__global__ void k_pack_dir(int ntru, int npair, int rb, int t, const double* filt64, float* out);
__global__ void k_pack_fwd4(int npair, int ncsp, const float* filt32, float* fwd4);
__global__ void k_twiddle(int n, double* tw64, float* tw32);

// Which inverse transform a launch performs.
enum InvKind : int {
  INV_SCALAR = 0,  // sp2fc with qi (or sp2fcdmu with qj) of a run of fields
  INV_UV = 1,      // dv2uv: (D, zeta) -> (u cos, v cos), terms qv and qu
};

// Device tables of one truncation (all FP64 values computed on the GPU).
struct Tables {
  int ntru = 0, nlat = 0, nlon = 0, npair = 0, ncsp = 0;
  int kbmax = 0;
  double* sid = nullptr;      // [nlat]
  double* gwd = nullptr;      // [nlat]
  double* tab = nullptr;      // [8][npair][ncsp] unfiltered
  double* filt64 = nullptr;   // [8][npair][ncsp] filter folded in
  float* filt32 = nullptr;
  float* fwd4 = nullptr;      // [8][npair/4][ncsp][4]: filt32 with 4 pairs side by side
  double* skgpsp = nullptr;   // [ntru+1]
  double* skspgp = nullptr;
  uint16_t* inv_bf[NTAB] = {};  // split-BF16 inverse table operands (3 planes)
  float* inv_f32[NTAB] = {};    // FP32-MFMA inverse table operands
  double* inv_f64[NTAB] = {};   // FP64-MFMA inverse table operands (same layout)
  long inv_elems = 0;           // elements of one inv_f32 (inv_f64) table
  float* dir_f32[NTAB] = {};    // MFMA direct table operands (qc, qe, qm, qq)
  int dir_rb = 0;               // row tiles of 16 per (m, parity)
  double* tw64 = nullptr;     // [nlon] complex
  float* tw32 = nullptr;
  double plavor = 0.0;
};

// Launch arguments of the fused/unfused inverse and direct transforms.
struct InvArgs {
  const void* sp;   // spectral input (member stride sp_stride reals)
  void* out;        // grid (fused) or Fourier (unfused) output
  long sp_stride, out_stride;
  int nmem;
  int kind;
  int table;        // INV_SCALAR: TQI or TQJ
  int src0, dst0, nfield;               // INV_SCALAR field runs
  int vor0, div0, u0, v0, nlev;         // INV_UV
  double plavor;
  int var;          // tile variant (0 default, 1 half the fields per block)
  int dlam;         // INV_SCALAR: times i m before the FFT (d/dlambda)
  // MODE_BF16X3: the parity-ordered spectral operand (launch_pack) of
  // fields [bpk_f0, ...), bpk_stride floats per member
  const void* bpk;
  long bpk_stride;
  int bpk_f0;
};

struct FwdArgs {
  const void* in;   // grid (fused) or Fourier (unfused) input
  void* sp;         // spectral output (or partial sums when npb > 1)
  long in_stride, sp_stride;
  int nmem;
  int src0, dst0, nfield;
  int table;        // TQC
  int var;          // tile variant (0 default, 1 half the fields per block)
};

// One term of a multi-input direct transform: out[oc] += sign * t * X_in[ic]
struct DirxTerm {
  int8_t in, ic, oc, sign;
  int8_t table;
};
// The term lists of the multi-input direct transforms (folded branches of
// src/legmod.f90): {input, input component, output component, sign, table}
// mktend (:538-549, fu <- gvz, fv <- guz as gridpointa calls it):
//   D = qq.ke - qe.fv + i qm.fu     inputs (ke, fv, fu)
//   T = qe.vt + qc.tn - i qm.ut     inputs (vt, tn, ut)
//   Z = qe.fu + i qm.fv             inputs (fu, fv)
// qtend (:468-496): q = qe.vq + qc.qn - i qm.uq   inputs (qn, uq, vq)
inline constexpr DirxTerm MKTEND_D[6] = {{0, 0, 0, +1, TQQ}, {1, 0, 0, -1, TQE},
                                         {2, 1, 0, -1, TQM}, {0, 1, 1, +1, TQQ},
                                         {1, 1, 1, -1, TQE}, {2, 0, 1, +1, TQM}};
inline constexpr DirxTerm MKTEND_T[6] = {{0, 0, 0, +1, TQE}, {1, 0, 0, +1, TQC},
                                         {2, 1, 0, +1, TQM}, {0, 1, 1, +1, TQE},
                                         {1, 1, 1, +1, TQC}, {2, 0, 1, -1, TQM}};
inline constexpr DirxTerm MKTEND_Z[4] = {{0, 0, 0, +1, TQE}, {1, 1, 0, -1, TQM},
                                         {0, 1, 1, +1, TQE}, {1, 0, 1, +1, TQM}};
inline constexpr DirxTerm QTEND_Q[6] = {{2, 0, 0, +1, TQE}, {0, 0, 0, +1, TQC},
                                        {1, 1, 0, +1, TQM}, {2, 1, 1, +1, TQE},
                                        {0, 1, 1, +1, TQC}, {1, 0, 1, -1, TQM}};

struct DirxArgs {
  const void* in[3];  // Fourier inputs fc(2, NLON/2, NLAT, nlev) per member
  void* out;          // spectral output (2*NCSP, nlev) per member
  int nin, nlev, nterm;
  DirxTerm terms[8];
  int nmem;                   // members (>= 1)
  long in_stride, out_stride; // member strides of the inputs / the output, in reals
};

int launch_dirx(const Tables& T, int mode, const DirxArgs& a, void* scratch, size_t sbytes,
                hipStream_t s);
size_t dirx_scratch_bytes(const Tables& T, int mode, int nlev);  // nlev: member-levels

int launch_inverse(const Tables& T, int mode, bool fused, const InvArgs& a, hipStream_t s);
// the sp2fc (aq), sp2fcdmu (aj) and dv2uv (au) transforms of a round trip,
// fused with fc2gp, in one launch; tile variant aq.var
int launch_inverse_rt(const Tables& T, int mode, const InvArgs& aq, const InvArgs& aj,
                      const InvArgs& au, hipStream_t s);
int launch_forward(const Tables& T, int mode, bool fused, const FwdArgs& a, void* scratch,
                   size_t scratch_bytes, hipStream_t s);
size_t forward_scratch_bytes(const Tables& T, int mode, int nmem, int nfield);
// Fourier rows of the T127 FP32 direct transform of grid input
size_t fwd_fft_scratch_bytes(const Tables& T, int nmem, int nfield);
// in-place FFTs of lot rows of n values, for nmem members mstride values
// apart; the forward FFT (dir -1) may read its rows from src instead
// (smstride per member) and write only the first kmax complex coefficients
// of each row (0: all n/2)
int launch_fft(const Tables& T, int mode, int dir, void* a, int n, long lot, hipStream_t s,
               int nmem = 1, long mstride = 0, const void* src = nullptr, long smstride = 0,
               int kmax = 0);
// parity-ordered copy (PK layout, sht_common.h) of fp32 spectral fields
// [f0, f0+nf) of nmem members for the split-BF16 inverse; out holds nmem *
// nf * pk_field(ntru) floats
int launch_pack(const Tables& T, const float* sp, long sp_stride, int f0, int nf, int nmem,
                float* out, hipStream_t s);
inline size_t pack_bytes(const Tables& T, int nmem, int nf) {
  return sizeof(float) * (size_t)nmem * nf * pk_field(T.ntru);
}

// the device tables of a plan (defined with the plan object)
const Tables& plan_tables(const exo_sht_plan_s* p);

}  // namespace exo
