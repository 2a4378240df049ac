/*
 * exo_out.h -- the spectral output format of PlaSim on the GPU (part of
 * libexosht.so).
 *
 * writesp (exoplasim/plasim/src/outmod.f90:141-169) writes a spectral field
 * as single-precision values zf = pf * spnorm * pscale, with poff added to
 * the first (m=0, n=0 real) value; spnorm is (-1)^m / sqrt(2) per complex
 * coefficient of zonal wavenumber m (src/plasim.f90:1638-1651), the
 * normalisation the post-processor's sp2gp undoes (exoplasim/pyfft.f90).
 */
#ifndef EXO_OUT_H
#define EXO_OUT_H
#ifdef __cplusplus
extern "C" {
#endif

/* nfield spectral fields of truncation ntru (pf, NRSP reals each, field
 * stride pf_stride, double for EXO_F64 else float) -> zf (float, NRSP each,
 * field stride zf_stride) with the scale pscale and the per-field offsets
 * poff (device array of nfield doubles, or NULL for none); arithmetic in
 * the type of pf, as the model's real kind.  Asynchronous on stream. */
int exo_sht_writesp(int mode, int ntru, int nfield, const void *pf, long pf_stride,
                    double pscale, const double *poff, float *zf, long zf_stride, void *stream);

#ifdef __cplusplus
}
#endif
#endif
