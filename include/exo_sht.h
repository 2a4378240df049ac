/*
 * exo_sht.h -- C ABI of the MI355X spectral-transform library (libexosht.so).
 *
 * Two layers:
 *
 * 1. Drop-in entry points with the argument meaning of the reference's
 *    legmod / fftmod / gaussmod subroutines (exoplasim/plasim/src/).  They
 *    take host arrays in the reference layout, run on the GPU and return when
 *    the result is back in the caller's arrays.  The _f64 / _f32 suffix is
 *    the caller's `real` kind (MOST_PREC 8 or 4).  The Fortran ISO_C_BINDING
 *    shim (exoplasim_amd/fortran/exo_legmod_capi.f90 and exo_transforms.f90) exports the original
 *    subroutine names on top of them.
 *
 * 2. A device-resident batched API on plans: many independent members
 *    transformed per launch with inputs and outputs already in device memory,
 *    enqueued on a caller stream.  This is what the throughput benchmark
 *    measures.
 *
 * Layouts (reference, Fortran column-major):
 *   spectral  sp[2*w + {0,1}]  (re, im), w = m-major, n inner, NRSP reals
 *   Fourier   fc[(lat*(NLON/2) + m)*2 + {0,1}], lat north -> south
 *   grid      gp[lat*NLON + j]  (the same memory as the Fourier array)
 * Every entry point returns EXO_OK (0) or a negative EXO_ERR_* code; the
 * Fortran shim turns a non-zero code into the model's abort path.
 */
#ifndef EXO_SHT_H
#define EXO_SHT_H
#ifdef __cplusplus
extern "C" {
#endif

#define EXO_OK 0
#define EXO_ERR_ARG (-1)
#define EXO_ERR_UNSUPPORTED (-2)
#define EXO_ERR_LAUNCH (-3)
#define EXO_ERR_SCRATCH (-4)
#define EXO_ERR_HIP (-5)
#define EXO_ERR_NOINIT (-6)

/* Legendre arithmetic of a transform call */
#define EXO_F64 0     /* FP64 vector FMA; FP64 data and FFT (the plain path) */
#define EXO_F32 1     /* FP32 vector FMA; FP32 data and FFT                 */
#define EXO_F32MFMA 2 /* v_mfma_f32_16x16x4_f32, FP32 data and FFT          */
#define EXO_BF16X3 3  /* split-BF16 x3 on v_mfma_f32_16x16x32_bf16, FP32 acc */

/* legmod table ids (src/legmod.f90:13-20) */
#define EXO_QI 0
#define EXO_QJ 1
#define EXO_QC 2
#define EXO_QE 3
#define EXO_QM 4
#define EXO_QQ 5
#define EXO_QU 6
#define EXO_QV 7

const char *exo_strerror(int code);
int exo_device_count(void);

/* ====================================================================== */
/* 1. drop-in entry points (host arrays, synchronous)                      */
/* ====================================================================== */

/* inigau (src/gaussmod.f90:5): Gaussian abscissas sin(phi) and weights,
 * north to south, computed in FP64 on the GPU. */
int exo_inigau(int klat, double *pz0, double *pzw);

/* legini (src/legmod.f90:30): builds the FP64 Legendre tables on the GPU
 * from the model's sid/gwd (src/plasimmod.f90:577-578) and makes them the
 * current tables of the drop-in entry points.  sid/gwd may be NULL (then
 * computed by inigau on the GPU); skgpsp/skspgp (NTRU+1 values each,
 * src/legmod.f90:21-22) may be NULL for no filter.  nlev is the model's
 * NLEV (used by dv2uv/uv2dv/mktend/qtend).  Supported truncations:
 * 21, 42, 63, 127 with nlat = 32, 64, 96, 192. */
int exo_legini(int ntru, int nlat, int nlev, const double *sid, const double *gwd,
               const double *skgpsp, const double *skspgp, double plavor);
/* copy of the current FP64 tables, tables[8][nlat][ncsp] in the order
 * qi qj qc qe qm qq qu qv (southern half by symmetry) */
int exo_legini_tables(double *tables);

/* sp2fc (src/legmod.f90:230): sp(2,NCSP) -> fc(2,NLON/2,NLAT), nfl
 * consecutive field-levels (NRSP / NLAT*NLON reals apart) */
int exo_sp2fc_f64(const double *sp, double *fc, int nfl);
int exo_sp2fc_f32(const float *sp, float *fc, int nfl);
/* sp2fcdmu (src/legmod.f90:262) */
int exo_sp2fcdmu_f64(const double *sp, double *fc, int nfl);
int exo_sp2fcdmu_f32(const float *sp, float *fc, int nfl);
/* fc2sp (src/legmod.f90:178), single-process folded branch */
int exo_fc2sp_f64(const double *fc, double *sp, int nfl);
int exo_fc2sp_f32(const float *fc, float *sp, int nfl);
/* dv2uv (src/legmod.f90:312): pd, pz (NRSP x NLEV), pu, pv (NHOR x NLEV);
 * pz is absolute vorticity, plavor is subtracted from its (m=0,n=1) real
 * coefficient (the caller's array is not modified) */
int exo_dv2uv_f64(const double *pd, const double *pz, double *pu, double *pv);
int exo_dv2uv_f32(const float *pd, const float *pz, float *pu, float *pv);
/* uv2dv (src/legmod.f90:356), folded branch: pu, pv (NHOR x NLEV) Fourier
 * coefficients of u*cos, v*cos -> pd, pz (NRSP x NLEV) */
int exo_uv2dv_f64(const double *pu, const double *pv, double *pd, double *pz);
int exo_uv2dv_f32(const float *pu, const float *pv, float *pd, float *pz);
/* qtend (src/legmod.f90:433), folded branch: q (NRSP x NLEV) from the
 * Fourier coefficients qn, uq, vq (NHOR x NLEV) */
int exo_qtend_f64(double *q, const double *qn, const double *uq, const double *vq);
int exo_qtend_f32(float *q, const float *qn, const float *uq, const float *vq);
/* mktend (src/legmod.f90:505), folded branch: spectral tendencies d, t, z
 * (NRSP x NLEV) from the Fourier coefficients tn, fu, fv, ke, ut, vt
 * (NHOR x NLEV); the model passes fu <- gvz, fv <- guz (src/plasim.f90:2478) */
int exo_mktend_f64(double *d, double *t, double *z, const double *tn, const double *fu,
                   const double *fv, const double *ke, const double *ut, const double *vt);
int exo_mktend_f32(float *d, float *t, float *z, const float *tn, const float *fu,
                   const float *fv, const float *ke, const float *ut, const float *vt);
/* gp2fc / fc2gp (src/fftmod.f90:29,64), in place over lot rows of length n */
int exo_gp2fc_f64(double *a, int n, int lot);
int exo_gp2fc_f32(float *a, int n, int lot);
int exo_fc2gp_f64(double *a, int n, int lot);
int exo_fc2gp_f32(float *a, int n, int lot);

/* ====================================================================== */
/* 2. device-resident batched API                                          */
/* ====================================================================== */

typedef struct exo_sht_plan_s *exo_sht_plan_t;

/* plan of one truncation; arguments as exo_legini */
int exo_sht_plan_create(exo_sht_plan_t *plan, int ntru, const double *sid, const double *gwd,
                        const double *skgpsp, const double *skspgp, double plavor);
int exo_sht_plan_destroy(exo_sht_plan_t plan);
/* dims = {ntru, nlat, nlon, ncsp} */
int exo_sht_plan_dims(exo_sht_plan_t plan, int *dims);
int exo_sht_plan_grid(exo_sht_plan_t plan, double *sid, double *gwd);
int exo_sht_plan_tables(exo_sht_plan_t plan, double *tables);

/* Batched inverse transform fused with fc2gp: for each of nmem members,
 * field-levels [src0, src0+nfield) of sp (member stride sp_stride reals) ->
 * grid fields [dst0, ...) of gp (member stride gp_stride reals).  table is
 * EXO_QI (sp2fc) or EXO_QJ (sp2fcdmu).  Device pointers, element type
 * double for EXO_F64 else float; asynchronous on stream (hipStream_t, NULL
 * = default stream). */
int exo_sht_inverse(exo_sht_plan_t plan, int mode, int table, int nmem, const void *sp,
                    long sp_stride, int src0, void *gp, long gp_stride, int dst0, int nfield,
                    void *stream);
/* Batched dv2uv fused with fc2gp: vorticity levels vor0.., divergence
 * levels div0.. -> u*cos levels u0.., v*cos levels v0.. */
int exo_sht_inverse_uv(exo_sht_plan_t plan, int mode, int nmem, const void *sp, long sp_stride,
                       int vor0, int div0, void *gp, long gp_stride, int u0, int v0, int nlev,
                       void *stream);
/* Batched gp2fc fused with fc2sp (folded): grid fields [src0, ...) ->
 * spectral field-levels [dst0, ...). */
int exo_sht_direct(exo_sht_plan_t plan, int mode, int nmem, const void *gp, long gp_stride,
                   int src0, void *sp, long sp_stride, int dst0, int nfield, void *stream);
/* One round trip of nmem members with nlev levels (the benchmark unit):
 * sp_in holds nfl = 3*nlev+1 field-levels per member (vorticity, divergence,
 * temperature levels, then ln ps); gp receives nfl + 1 + 2*nlev grid fields
 * per member (the nfl fields, d(ln ps)/dmu, u*cos, v*cos); sp_out receives
 * the nfl spectral field-levels of gp2fc+fc2sp of the first nfl grid
 * fields. */
int exo_sht_roundtrip(exo_sht_plan_t plan, int mode, int nmem, int nlev, const void *sp_in,
                      void *gp, void *sp_out, void *stream);
/* The inverse half of exo_sht_roundtrip: the grid-point fields of the
 * dynamics step (invlegd and invlega, src/legmod.f90:635-677) -- sp2fc of the
 * nfl = 3*nlev+1 field-levels, sp2fcdmu of ln ps and dv2uv of the nlev
 * levels, each fused with fc2gp, in one kernel launch; same layouts. */
int exo_sht_inverse_all(exo_sht_plan_t plan, int mode, int nmem, int nlev, const void *sp_in,
                        void *gp, void *stream);

/* No counterpart in the reference: exo_sht_roundtrip for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, layouts and
 * results as exo_sht_roundtrip; it calls no hipMalloc or hipFree.  Its
 * scratch (the pre-split operand of EXO_BF16X3 and the partial direct sums
 * of T63 in FP64 and of T127) is a file-scope device array of
 * EXO_NOALLOC_SCRATCH_BYTES per GPU, held from the library's loading on
 * that GPU until the process ends; members run in chunks whose scratch
 * fits it.  EXO_ERR_SCRATCH when one member's scratch exceeds it (T127
 * L40 in EXO_F64 needs 96 MB per member).  The plan must be used on the
 * device it was created on.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other plans, other
 * threads) race and must be ordered by the caller. */
#define EXO_NOALLOC_SCRATCH_BYTES (64u << 20)
int exo_sht_roundtrip_noalloc(exo_sht_plan_t plan, int mode, int nmem, int nlev,
                              const void *sp_in, void *gp, void *sp_out, void *stream);

/* Device-resident forms of the legmod / fftmod routines that work in
 * Fourier space, on device arrays of nmem members (member strides in reals),
 * asynchronous on stream; element type double for EXO_F64 else float.
 * Fourier arrays are fc(2, NLON/2, NLAT) per field-level, the same memory as
 * a grid field.
 *   exo_sht_sp2fc   sp2fc / sp2fcdmu (src/legmod.f90:230,262), table EXO_QI
 *                   or EXO_QJ: spectral field-levels [src0, src0+nfield) ->
 *                   Fourier fields [dst0, ...)
 *   exo_sht_dv2uv   dv2uv (src/legmod.f90:312): vorticity levels vor0..,
 *                   divergence levels div0.. -> Fourier u*cos u0.., v*cos v0..
 *   exo_sht_fc2sp   fc2sp (src/legmod.f90:178), folded branch
 *   exo_sht_fft     gp2fc (dir -1) / fc2gp (dir +1) (src/fftmod.f90:29,64)
 *                   in place over lot rows of the plan's NLON
 *   exo_sht_uv2dv   uv2dv (src/legmod.f90:356): Fourier u, v (nlev levels per
 *                   member, member stride f_stride) -> spectral D, zeta
 *                   (member stride s_stride)
 *   exo_sht_qtend   qtend (src/legmod.f90:433)
 *   exo_sht_mktend  mktend (src/legmod.f90:505); the model passes fu <- gvz,
 *                   fv <- guz (src/plasim.f90:2478)
 * The multi-input direct transforms (uv2dv, qtend, mktend) sum on the vector
 * ALUs in FP64 (EXO_F64) or FP32 (every other mode). */
int exo_sht_sp2fc(exo_sht_plan_t plan, int mode, int table, int nmem, const void *sp,
                  long sp_stride, int src0, void *fc, long fc_stride, int dst0, int nfield,
                  void *stream);
int exo_sht_dv2uv(exo_sht_plan_t plan, int mode, int nmem, const void *sp, long sp_stride,
                  int vor0, int div0, void *fc, long fc_stride, int u0, int v0, int nlev,
                  void *stream);
int exo_sht_fc2sp(exo_sht_plan_t plan, int mode, int nmem, const void *fc, long fc_stride,
                  int src0, void *sp, long sp_stride, int dst0, int nfield, void *stream);
int exo_sht_fft(exo_sht_plan_t plan, int mode, int dir, void *a, long lot, void *stream);
int exo_sht_uv2dv(exo_sht_plan_t plan, int mode, int nmem, int nlev, const void *pu,
                  const void *pv, long f_stride, void *pd, void *pz, long s_stride, void *stream);
int exo_sht_qtend(exo_sht_plan_t plan, int mode, int nmem, int nlev, void *q, long s_stride,
                  const void *qn, const void *uq, const void *vq, long f_stride, void *stream);
int exo_sht_mktend(exo_sht_plan_t plan, int mode, int nmem, int nlev, void *d, void *t, void *z,
                   long s_stride, const void *tn, const void *fu, const void *fv, const void *ke,
                   const void *ut, const void *vt, long f_stride, void *stream);

/* Tile variants of the fused kernels: vars[0] the inverse sp2fc/sp2fcdmu
 * kernel and the one-launch inverse of exo_sht_inverse_all /
 * exo_sht_roundtrip, vars[1] the dv2uv kernel of exo_sht_inverse_uv,
 * vars[2] the direct kernel; 0 = default tiles, 1 = half the field-levels
 * per block (half the LDS, twice the blocks in flight), 2 = default tiles
 * on 512-thread blocks (twice the waves per block), 3 = half the
 * field-levels on 512-thread blocks.  exo_sht_autotune
 * times each variant (two untimed calls, then the average of eight
 * back-to-back calls) on a round trip of nmem members x nlev levels, keeps
 * the fastest of each for this mode and returns them in vars[3] (may be
 * NULL).
 * Calls on one plan share its scratch buffers: calls that overlap on
 * different streams race. */
int exo_sht_set_tiles(exo_sht_plan_t plan, int mode, int inv_var, int uv_var, int fwd_var);
int exo_sht_get_tiles(exo_sht_plan_t plan, int mode, int *vars);
int exo_sht_autotune(exo_sht_plan_t plan, int mode, int nmem, int nlev, void *stream, int *vars);

/* ====================================================================== */
/* 3. mixed-truncation ensembles                                           */
/* ====================================================================== */
/* No counterpart in the reference, which runs one model per process: an
 * ensemble of members with different truncations, each group in the layout
 * of exo_sht_roundtrip (sp_in and sp_out (nmem, 3*nlev+1, 2*NCSP), gp
 * (nmem, 5*nlev+2, NLAT, NLON), float device arrays).  Truncations 21, 42
 * and 63; modes EXO_F32, EXO_F32MFMA and EXO_BF16X3.  The round trips of all
 * members run as two persistent kernels (all inverse work items, then all
 * direct work items) captured in one hipGraph at create time, or, after
 * exo_sht_ens_set_schedule or autotune, as one batched round trip per group
 * in that graph; the arrays must stay allocated while the ensemble exists. */
typedef struct {
  int ntru, nlev, nmem;
  const void *sp_in;
  void *gp;
  void *sp_out;
} exo_sht_ens_group;
typedef struct exo_sht_ens_s *exo_sht_ens_t;

int exo_sht_ens_create(exo_sht_ens_t *ens, int mode, int ngroup, const exo_sht_ens_group *groups);
int exo_sht_ens_destroy(exo_sht_ens_t ens);
/* one launch of the captured graph on stream */
int exo_sht_ens_run(exo_sht_ens_t ens, void *stream);
/* the same kernels launched directly (no graph) */
int exo_sht_ens_run_eager(exo_sht_ens_t ens, void *stream);
/* the same round trips as one batched exo_sht_roundtrip per group */
int exo_sht_ens_run_groups(exo_sht_ens_t ens, void *stream);
/* the same round trips one member at a time (4 launches per member) */
int exo_sht_ens_run_members(exo_sht_ens_t ens, void *stream);
/* tile variants per group, vars[2*k] inverse and vars[2*k+1] direct of
 * group k, 0-3 as for exo_sht_set_tiles.  Variants 2 and 3 run the persistent
 * kernel with 512-thread blocks, so they hold for every group of that
 * kernel or for none (else EXO_ERR_ARG).  set_tiles captures the graph
 * again.  autotune times every combination per truncation (0/1, and 2/3),
 * keeps the fastest and returns it in vars (may be NULL). */
int exo_sht_ens_get_tiles(exo_sht_ens_t ens, int *vars);
int exo_sht_ens_set_tiles(exo_sht_ens_t ens, const int *vars);
int exo_sht_ens_autotune(exo_sht_ens_t ens, void *stream, int *vars);
/* What the graph of exo_sht_ens_run holds (and exo_sht_ens_run_eager
 * launches): the two persistent kernels (EXO_ENS_PERSISTENT, the default)
 * or one batched exo_sht_roundtrip per group (EXO_ENS_BATCHED, the kernels
 * of exo_sht_ens_run_groups).  set_schedule captures the graph again (for
 * EXO_ENS_BATCHED after one direct pass, which writes the outputs);
 * autotune, after choosing the persistent kernels' tile variants, lets each
 * group's plan choose those of its batched round trip (exo_sht_autotune),
 * times both schedules and keeps the faster.  set_tiles sets the variants
 * of both schedules. */
#define EXO_ENS_PERSISTENT 0
#define EXO_ENS_BATCHED 1
int exo_sht_ens_get_schedule(exo_sht_ens_t ens, int *sched);
int exo_sht_ens_set_schedule(exo_sht_ens_t ens, int sched);

#ifdef __cplusplus
}
#endif
#endif
