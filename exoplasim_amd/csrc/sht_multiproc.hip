// sht_multiproc.hip -- what PlaSim's multi-process (NPRO > 1) path computes
// around the transforms, on one GPU (declared in include/exo_mp.h):
//
//   k_fc2sp_part  the unfolded direct Legendre transform of one process's
//                 latitudes (src/legmod.f90:191-202): partial spectral sums
//                 over the latitudes [lat0, lat0 + nlpp)
//   k_sumsc       mpsumsc (src/mpimod.f90:279-291): the sum over processes
//                 of those partial sums, each process receiving its NSPP
//                 slice (MPI_Reduce_scatter with counts nscatsp = NSPP,
//                 src/plasimmod.f90:306)
//   k_trunc       mrtrunc / mrexpand (src/mpimod_multi.f90:117-153): copy
//                 spectral fields between two triangular truncations
//
// With NPRO processes a latitude band is contiguous: process r holds the
// Gaussian latitudes r NLPP .. (r+1) NLPP - 1, north to south, as
// mpscdn/mpscgp scatter them (src/mpimod.f90:116-143).  Each kernel reads
// and writes its arrays once, coalesced: HBM-bound.
#include "../../include/exo_mp.h"
#include "sht_plan.h"

namespace exo {

// one thread per (member, field, spectral real of NESP); the table of a
// southern latitude is its northern partner's with the parity sign
template <typename R>
__global__ void k_fc2sp_part(const R* fc, long fc_stride, int src0, R* sp, long sp_stride,
                             int dst0, int nfield, int nesp, int lat0, int nlpp, const R* qc,
                             int ntru, int nlat, int nlon, long total) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int x = (int)(i % nesp);
  const long fm = i / nesp;
  const int f = (int)(fm % nfield), mem = (int)(fm / nfield);
  R* out = sp + mem * sp_stride + (long)(dst0 + f) * nesp + x;
  const int ncsp = (ntru + 1) * (ntru + 2) / 2;
  const int w = x >> 1, ri = x & 1;
  if (w >= ncsp) {
    *out = R(0);  // padding of NESP past NRSP
    return;
  }
  const int m = m_of_w(ntru, w);
  const int n = m + (w - w0_of(ntru, m));
  const bool odd = ((n - m) & 1) != 0;
  const int npair = nlat / 2;
  const R* src = fc + mem * fc_stride + (long)(src0 + f) * nlat * nlon + 2 * m + ri;
  R acc = R(0);
  for (int l = lat0; l < lat0 + nlpp; ++l) {
    const bool south = l >= npair;
    const int p = south ? nlat - 1 - l : l;
    R q = qc[(long)p * ncsp + w];
    if (south && odd) q = -q;
    acc += q * src[(long)l * nlon];
  }
  *out = acc;
}

// psp[r][f][i] = sum over r' (ascending) of psf[r'][f][r NSPP + i]
template <typename R>
__global__ void k_sumsc(const R* psf, R* psp, int nproc, int nfield, int nspp, long total) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int k = (int)(i % nspp);
  const long rf = i / nspp;
  const int f = (int)(rf % nfield), r = (int)(rf / nfield);
  const long nesp = (long)nspp * nproc;
  R s = psf[(long)f * nesp + (long)r * nspp + k];
  for (int q = 1; q < nproc; ++q) s += psf[((long)q * nfield + f) * nesp + (long)r * nspp + k];
  psp[i] = s;
}

// mrtrunc (expand = 0): lo(m, n) = hi(m, n) for n <= ntlo; mrexpand
// (expand = 1): hi(m, n) = lo(m, n) for n <= ntlo, and the reference's
// clearing of hi before it, which covers the first nthi complex values only
// (src/mpimod_multi.f90:141); one thread per complex value of the target
template <typename R>
__global__ void k_trunc(const R* src, long src_stride, R* dst, long dst_stride, int nthi,
                        int ntlo, int expand, int nfield, long total) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int nt_dst = expand ? nthi : ntlo, nt_src = expand ? ntlo : nthi;
  const int ncsp_dst = (nt_dst + 1) * (nt_dst + 2) / 2;
  const int w = (int)(i % ncsp_dst);
  const int f = (int)(i / ncsp_dst);
  const int m = m_of_w(nt_dst, w);
  const int n = m + (w - w0_of(nt_dst, m));
  R* d = dst + f * dst_stride + 2 * w;
  if (m <= ntlo && n <= ntlo) {
    const R* s = src + f * src_stride + 2 * (w0_of(nt_src, m) + (n - m));
    d[0] = s[0];
    d[1] = s[1];
  } else if (expand && w < nthi) {
    d[0] = R(0);
    d[1] = R(0);
  }
}

}  // namespace exo

using namespace exo;

static bool mode_valid(int mode) { return mode >= EXO_F64 && mode <= EXO_BF16X3; }

extern "C" int exo_sht_fc2sp_part(exo_sht_plan_t p, int mode, int nmem, const void* fc,
                                  long fc_stride, int src0, int lat0, int nlpp, void* sp,
                                  long sp_stride, int dst0, int nesp, int nfield, void* stream) {
  if (!p || !mode_valid(mode) || nmem < 0 || nfield < 0 || src0 < 0 || dst0 < 0)
    return EXO_ERR_ARG;
  const Tables& tb = p->tb;
  if (lat0 < 0 || nlpp < 1 || lat0 + nlpp > tb.nlat || nesp < 2 * tb.ncsp)
    return EXO_ERR_ARG;
  if (nmem == 0 || nfield == 0) return EXO_OK;
  if (!fc || !sp) return EXO_ERR_ARG;
  const long total = (long)nmem * nfield * nesp;
  const unsigned nb = (unsigned)((total + 255) / 256);
  const long tsz = (long)tb.npair * tb.ncsp;
  hipStream_t s = (hipStream_t)stream;
  if (mode == EXO_F64)
    k_fc2sp_part<double><<<nb, 256, 0, s>>>((const double*)fc, fc_stride, src0, (double*)sp,
                                            sp_stride, dst0, nfield, nesp, lat0, nlpp,
                                            tb.filt64 + TQC * tsz, tb.ntru, tb.nlat, tb.nlon,
                                            total);
  else
    k_fc2sp_part<float><<<nb, 256, 0, s>>>((const float*)fc, fc_stride, src0, (float*)sp,
                                           sp_stride, dst0, nfield, nesp, lat0, nlpp,
                                           tb.filt32 + TQC * tsz, tb.ntru, tb.nlat, tb.nlon,
                                           total);
  return hipGetLastError() == hipSuccess ? EXO_OK : EXO_ERR_LAUNCH;
}

extern "C" int exo_sht_sumsc(int mode, int nproc, int nspp, int nfield, const void* psf,
                             void* psp, void* stream) {
  if (!mode_valid(mode) || nproc < 1 || nspp < 1 || nfield < 0) return EXO_ERR_ARG;
  if (nfield == 0) return EXO_OK;
  if (!psf || !psp) return EXO_ERR_ARG;
  const long total = (long)nproc * nfield * nspp;
  const unsigned nb = (unsigned)((total + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  if (mode == EXO_F64)
    k_sumsc<double><<<nb, 256, 0, s>>>((const double*)psf, (double*)psp, nproc, nfield, nspp,
                                       total);
  else
    k_sumsc<float><<<nb, 256, 0, s>>>((const float*)psf, (float*)psp, nproc, nfield, nspp,
                                      total);
  return hipGetLastError() == hipSuccess ? EXO_OK : EXO_ERR_LAUNCH;
}

static int trunc_impl(int mode, int nthi, int ntlo, int expand, int nfield, const void* src,
                      long src_stride, void* dst, long dst_stride, hipStream_t s) {
  if (!mode_valid(mode) || ntlo < 0 || nthi < ntlo || nthi > 1023 || nfield < 0)
    return EXO_ERR_ARG;
  if (nfield == 0) return EXO_OK;
  if (!src || !dst) return EXO_ERR_ARG;
  const int ntd = expand ? nthi : ntlo;
  const long total = (long)nfield * ((ntd + 1) * (ntd + 2) / 2);
  const unsigned nb = (unsigned)((total + 255) / 256);
  if (mode == EXO_F64)
    k_trunc<double><<<nb, 256, 0, s>>>((const double*)src, src_stride, (double*)dst, dst_stride,
                                       nthi, ntlo, expand, nfield, total);
  else
    k_trunc<float><<<nb, 256, 0, s>>>((const float*)src, src_stride, (float*)dst, dst_stride,
                                      nthi, ntlo, expand, nfield, total);
  return hipGetLastError() == hipSuccess ? EXO_OK : EXO_ERR_LAUNCH;
}

extern "C" int exo_sht_mrtrunc(int mode, int nthi, int ntlo, int nfield, const void* sphi,
                               long hi_stride, void* splo, long lo_stride, void* stream) {
  return trunc_impl(mode, nthi, ntlo, 0, nfield, sphi, hi_stride, splo, lo_stride,
                    (hipStream_t)stream);
}

extern "C" int exo_sht_mrexpand(int mode, int ntlo, int nthi, int nfield, const void* splo,
                                long lo_stride, void* sphi, long hi_stride, void* stream) {
  return trunc_impl(mode, nthi, ntlo, 1, nfield, splo, lo_stride, sphi, hi_stride,
                    (hipStream_t)stream);
}
