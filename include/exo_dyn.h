/*
 * exo_dyn.h -- C ABI of the grid-point dynamics between the inverse and the
 * direct transforms (part of libexosht.so): calcgp, the Fourier-space
 * d/dlambda of ln(ps) and the device-resident gridpointa transform group.
 *
 * Reference: exoplasim/plasim/src/plasim.f90
 *   gridpointa  :2360-2585  (invlega, gpmt, fc2gp, exp, calcgp, products,
 *                            gp2fc, fc2sp, mktend, qtend, mpsumsc)
 *   calcgp      :2587-2750
 *   initsi      :1709-1770  (the vertical set-up calcgp reads)
 * Fields are the model's grid-point arrays: (NHOR, NLEV) column-major, one
 * level of NHOR = NLON * NLAT values after another.
 */
#ifndef EXO_DYN_H
#define EXO_DYN_H
#include "exo_sht.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct exo_dyn_s *exo_dyn_t;

/* Vertical set-up of the grid-point dynamics, copied to the device in FP64
 * and FP32: sigmah, dsigma, rdsig, t0, t01s2, tkp (nlev values each; t0 and
 * its differences already divided by ct as the model holds them), c and g
 * (nlev x nlev, column-major: c(i,j) = c[(j-1)*nlev + i-1]) from initsi, and
 * the scalars akap, rdbrv, adv (the model's ADV), psurf, ww.  2 <= nlev <= 64. */
int exo_dyn_create(exo_dyn_t *dyn, int nlev, const double *sigmah, const double *dsigma,
                   const double *rdsig, const double *t0, const double *t01s2, const double *tkp,
                   const double *c, const double *g, double akap, double rdbrv, double adv,
                   double psurf, double ww);
int exo_dyn_destroy(exo_dyn_t dyn);

/* Device pointers of one member's calcgp fields; (nlev) fields hold nlev
 * levels of nhor values.  gut .. gvq may be NULL: when given they receive
 * gridpointa's products gu*gt, gv*gt, gu^2 + gv^2 + gphi, and (nqspec = 1)
 * gu*gq, gv*gq (src/plasim.f90:2436-2446). */
typedef struct {
  const void *gu, *gv, *gt, *gd, *gz, *gq; /* (nlev) in                       */
  const void *dq;                          /* (nlev) in, read when nqspec = 0 */
  void *gp;                                /* in: ln ps (exp_gp) or ps/psurf;
                                              out: ps/psurf                   */
  const void *gpj, *gpm, *rcsq;            /* in: d ln ps/dmu, d ln ps/dlambda
                                              (gpmt), 1/cos^2 per point       */
  void *gtn, *gqn, *guz, *gvz, *gphi, *dw; /* (nlev) out                      */
  void *gvp;                               /* out                             */
  void *gut, *gvt, *gke, *guq, *gvq;       /* (nlev) out, optional            */
} exo_calcgp_fields;

/* calcgp (src/plasim.f90:2587) for nmem members whose fields all lie
 * member_stride elements apart; exp_gp != 0 first replaces gp by exp(gp)
 * (src/plasim.f90:2434).  Element type double for EXO_F64, float for the
 * other modes (and arithmetic in that type); asynchronous on stream. */
int exo_dyn_calcgp(exo_dyn_t dyn, int mode, int nmem, long member_stride, int nhor, int nqspec,
                   int exp_gp, const exo_calcgp_fields *f, void *stream);

/* gpmt (src/plasim.f90:2425-2430): the Fourier coefficients of d/dlambda,
 * out(2m) = -m in(2m+1), out(2m+1) = m in(2m) per row of nlon values, for
 * nrow rows; in and out device arrays of the mode's type (may alias). */
int exo_dyn_gpmt(int mode, int nlon, long nrow, const void *in, void *out, void *stream);

/* The dynamics' transform group of one time step on the device, gridpointa
 * (src/plasim.f90:2360-2585) without its diagnostics: invlega (sp2fc of sd
 * st sz [sq] sp, sp2fcdmu of sp, dv2uv) fused with fc2gp, gpmt, gp = exp(gp),
 * calcgp and the products, gp2fc, fc2sp(gvpp), mktend and [qtend].  Arrays
 * of nmem members, all on the device, element type as exo_dyn_calcgp:
 *   sp_state  per member 4*nlev+1 spectral field-levels of NRSP reals:
 *             sz (nlev, absolute vorticity), sd, st, sq (nlev each), sp
 *   sp_tend   the same layout: szt, sdt, stt, sqt (written when nqspec = 1;
 *             left as they are otherwise), spt; what mpsumsc hands on with
 *             one process (src/mpimod_stub.f90:91-97)
 *   work      exo_gridpointa_fields(nlev) grid fields of NHOR values:
 *               0            gz gd gt gq (nlev each), then gp gpj gpm
 *               4 nlev + 3   gu gv (nlev each)
 *               6 nlev + 3   dq (nlev): the caller's grid-point humidity,
 *                            read when nqspec = 0 (gq is then the caller's
 *                            too, as the model's)
 *               7 nlev + 3   rcsq (written here from the plan's latitudes)
 *               7 nlev + 4   gtn gqn guz gvz gut gvt gke guq gvq (nlev each),
 *                            gvp: Fourier coefficients on return
 *              16 nlev + 5   gphi dw (nlev each), grid values on return
 * The plan's tables, filters and plavor and the dyn's vertical set-up are
 * used; strides are in elements. */
int exo_gridpointa_fields(int nlev);
int exo_gridpointa(exo_sht_plan_t plan, exo_dyn_t dyn, int mode, int nmem, int nqspec,
                   const void *sp_state, long sp_stride, void *work, long work_stride,
                   void *sp_tend, long tend_stride, void *stream);

#ifdef __cplusplus
}
#endif
#endif
