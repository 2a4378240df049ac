// sht_output.hip -- writesp's normalisation of the spectral output
// (src/outmod.f90:141-169, spnorm of src/plasim.f90:1638-1651) on the GPU,
// declared in include/exo_out.h.  One thread per spectral real: read once,
// written once as float.
#include "../../include/exo_out.h"
#include "../../include/exo_sht.h"
#include "sht_internal.h"

namespace exo {

template <typename R>
__global__ void k_writesp(int ntru, int nrsp, int nfield, const R* pf, long pf_stride, R pscale,
                          const double* poff, float* zf, long zf_stride) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)nrsp * nfield) return;
  const int x = (int)(i % nrsp), f = (int)(i / nrsp);
  const int m = m_of_w(ntru, x >> 1);
  // zrsq2 = 1.0 / sqrt(2.0) in the model's real kind, its sign flipped
  // for every odd m
  R norm = R(1) / sqrt(R(2));
  if (m & 1) norm = -norm;
  float z = (float)(pf[f * pf_stride + x] * norm * pscale);
  if (x == 0 && poff) z = (float)((R)z + (R)poff[f]);
  zf[f * zf_stride + x] = z;
}

}  // namespace exo

using namespace exo;

extern "C" int exo_sht_writesp(int mode, int ntru, int nfield, const void* pf, long pf_stride,
                               double pscale, const double* poff, float* zf, long zf_stride,
                               void* stream) {
  if (mode < EXO_F64 || mode > EXO_BF16X3 || ntru < 0 || ntru > 1023 || nfield < 0)
    return EXO_ERR_ARG;
  if (nfield == 0) return EXO_OK;
  if (!pf || !zf) return EXO_ERR_ARG;
  const int nrsp = (ntru + 1) * (ntru + 2);
  const long total = (long)nrsp * nfield;
  const unsigned nb = (unsigned)((total + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  if (mode == EXO_F64)
    k_writesp<double><<<nb, 256, 0, s>>>(ntru, nrsp, nfield, (const double*)pf, pf_stride,
                                         pscale, poff, zf, zf_stride);
  else
    k_writesp<float><<<nb, 256, 0, s>>>(ntru, nrsp, nfield, (const float*)pf, pf_stride,
                                        (float)pscale, poff, zf, zf_stride);
  return hipGetLastError() == hipSuccess ? EXO_OK : EXO_ERR_LAUNCH;
}
