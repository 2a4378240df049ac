/*
 * exo_si.h -- C ABI of the semi-implicit spectral solve on the GPU (FP64),
 * part of libexosht.so.
 *
 * The reference builds bm1 = inverse(delt^2 (t0 dsigma + g.tau) + I/(n(n+1)))
 * for n = 1..NTRU in makebm (exoplasim/plasim/src/plasim.f90:1683-1703, with
 * minvers/ludcmp/lubksb :1804-1888) and applies it in spectrala
 * (src/plasim.f90:2817-2876): the divergence at time t from eq. 17 of
 * Hoskins & Simmons (1975), then the ln(ps) and temperature updates.
 * Matrices are column-major NLEV x NLEV (g(NLEV,NLEV), tau(NLEV,NLEV)),
 * spectral arrays (NSP, NLEV) column-major with NSP = (NTRU+1)(NTRU+2) reals
 * (one process, NSPP = NRSP).
 */
#ifndef EXO_SI_H
#define EXO_SI_H
#ifdef __cplusplus
extern "C" {
#endif

#define EXO_SI_OK 0
#define EXO_SI_ERR_ARG (-1)
#define EXO_SI_ERR_HIP (-5)
#define EXO_SI_ERR_SINGULAR (-7)

typedef struct exo_si_s *exo_si_t;

/* copies the vertical-scheme matrices (from initsi, src/plasim.f90:1709)
 * to the device; nlev <= 64 */
int exo_si_create(exo_si_t *si, int nlev, int ntru, const double *t0, const double *dsigma,
                  const double *g, const double *tau);
int exo_si_destroy(exo_si_t si);
/* makebm (src/plasim.f90:1683) on the GPU for time step delt */
int exo_si_makebm(exo_si_t si, double delt, void *stream);
/* host copy of bm1(NLEV, NLEV, NTRU) */
int exo_si_bm1(exo_si_t si, double *bm1);
/* section 1-3 of spectrala's semi-implicit step on device arrays of the
 * model's real kind (fp64 != 0: double, else float), FP64 arithmetic:
 *   in : sop(NSP) surface geopotential, spm(NSP) ln ps at t-dt,
 *        atm(NSP,NLEV) temperature at t-dt, adm(NSP,NLEV) divergence at t-dt
 *   in/out: sdt (divergence tendency -> divergence at t), spt (-script P ->
 *        updated), stt (script T -> updated) */
int exo_si_apply(exo_si_t si, int fp64, double delt, const void *sop, const void *spm,
                 const void *atm, const void *adm, void *stt, void *sdt, void *spt,
                 void *stream);

#ifdef __cplusplus
}
#endif
#endif
