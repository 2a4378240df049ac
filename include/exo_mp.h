/*
 * exo_mp.h -- what PlaSim's multi-process path (NPRO > 1) computes around
 * the transforms, on one GPU (part of libexosht.so).
 *
 * With NPRO processes each holds NLPP = NLAT / NPRO consecutive Gaussian
 * latitudes (process r: r*NLPP .. (r+1)*NLPP - 1, north to south, as
 * mpscdn / mpscgp scatter them, src/mpimod.f90:116-143) and NSPP =
 * ceil(NRSP / NPRO) spectral reals; spectral arrays have NESP = NSPP * NPRO
 * reals (src/plasimmod.f90:102-113).  Device arrays of the mode's type
 * (double for EXO_F64, float otherwise), strides in elements, asynchronous
 * on stream.
 */
#ifndef EXO_MP_H
#define EXO_MP_H
#include "exo_sht.h"
#ifdef __cplusplus
extern "C" {
#endif

/* fc2sp, multi-process branch (src/legmod.f90:191-202): the unfolded partial
 * sums over the latitudes [lat0, lat0 + nlpp) of Fourier fields
 * fc(2, NLON/2, NLAT) (the whole grid; only those latitudes are read) into
 * spectral partials of nesp >= NRSP reals per field, zero past NRSP.  The
 * plan's GP->SP filter applies as in the reference. */
int exo_sht_fc2sp_part(exo_sht_plan_t plan, int mode, int nmem, const void *fc, long fc_stride,
                       int src0, int lat0, int nlpp, void *sp, long sp_stride, int dst0,
                       int nesp, int nfield, void *stream);

/* mpsumsc (src/mpimod.f90:279-291): psf holds the nproc processes' partial
 * sums [nproc][nfield][NESP] (NESP = nspp * nproc); psp receives
 * [nproc][nfield][nspp], process r's slice r*nspp .. of the sum over the
 * processes (summed in process order). */
int exo_sht_sumsc(int mode, int nproc, int nspp, int nfield, const void *psf, void *psp,
                  void *stream);

/* mrtrunc / mrexpand (src/mpimod_multi.f90:117-153): spectral fields from
 * truncation nthi to ntlo <= nthi and back, in the reference's (m-major)
 * order.  mrexpand clears the first nthi complex values of each target
 * field before it copies, as the reference does (:141); the rest of the
 * target keeps its contents. */
int exo_sht_mrtrunc(int mode, int nthi, int ntlo, int nfield, const void *sphi, long hi_stride,
                    void *splo, long lo_stride, void *stream);
int exo_sht_mrexpand(int mode, int ntlo, int nthi, int nfield, const void *splo,
                     long lo_stride, void *sphi, long hi_stride, void *stream);

#ifdef __cplusplus
}
#endif
#endif
