/*
 * exo_oracle.h -- FP64 CPU restatement of PlaSim's spectral-transform core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the test oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (exoplasim_amd) never links or calls it.
 *
 * Array conventions follow the reference (column-major Fortran, first index
 * fastest):
 *   spectral  sp[2*w + {0,1}]                 w = 0..ncsp-1, m-major, n inner
 *   Fourier   fc[(l*(nlon/2) + m)*2 + {0,1}]  l = 0..nlat-1 north->south
 *   grid      gp[l*nlon + j]                  (the same memory as fc)
 *   tables    q[l*ncsp + w]                   (qi(NCSP,NLPP) in legmod.f90:13)
 * All arithmetic is IEEE double with the reference's evaluation order.
 */
#ifndef EXO_ORACLE_H
#define EXO_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- Gaussian grid (src/gaussmod.f90:5-86) ---- */
double ora_ql(int k, double p);
double ora_qld(int k, double p);
void ora_inigau(int klat, double *pz0, double *pzw);

/* ---- Legendre tables (src/legmod.f90:30-116) ----
 * tables = 8 consecutive blocks of ncsp*nlat doubles in the order
 * qi, qj, qc, qe, qm, qq, qu, qv. */
enum { ORA_QI = 0, ORA_QJ, ORA_QC, ORA_QE, ORA_QM, ORA_QQ, ORA_QU, ORA_QV };
void ora_legini(int ntru, int nlat, const double *sid, const double *gwd,
                double *tables);
/* physics filters (src/legmod.f90:123-155); sk[i] belongs to total
 * wavenumber n=i (Fortran index i+1). */
void ora_filters(int ntru, int nfilter, int ngptfilter, int nspvfilter,
                 double filterkappa, int nfilterexp, double landhoskn0,
                 double *skgpsp, double *skspgp);

/* ---- Legendre transforms (src/legmod.f90) ---- */
void ora_sp2fc(int ntru, int nlat, int nlon, const double *qi,
               const double *skspgp, const double *sp, double *fc);
void ora_fc2sp(int ntru, int nlat, int nlon, const double *qc,
               const double *skgpsp, const double *fc, double *sp);
void ora_fc2sp_unfolded(int ntru, int nlpp, int nlon, const double *qc,
                        const double *skgpsp, const double *fc, double *sp);
void ora_dv2uv(int ntru, int nlat, int nlon, int nlev, const double *qu,
               const double *qv, const double *skspgp, double plavor,
               const double *pd, const double *pz, double *pu, double *pv);
void ora_uv2dv(int ntru, int nlat, int nlon, int nlev, const double *qe,
               const double *qm, const double *skgpsp, const double *pu,
               const double *pv, double *pd, double *pz);
void ora_qtend(int ntru, int nlat, int nlon, int nlev, const double *qc,
               const double *qe, const double *qm, const double *skgpsp,
               const double *qn, const double *uq, const double *vq,
               double *q);
void ora_mktend(int ntru, int nlat, int nlon, int nlev, const double *qc,
                const double *qe, const double *qm, const double *qq,
                const double *skgpsp, const double *tn, const double *fu,
                const double *fv, const double *ke, const double *ut,
                const double *vt, double *d, double *t, double *z);

/* ---- real FFT of fftmod (src/fftmod.f90) ----
 * returns 0, or -1 if n is not one of fftmod's supported lengths */
int ora_fft_supported(int n);
int ora_gp2fc(double *a, int n, int lot);
int ora_fc2gp(double *a, int n, int lot);

/* ---- real FFT of fft991mod by its documented definition
 * (src/fft991mod.f90:148-155); any even n ---- */
void ora_gp2fc_dft(double *a, int n, int lot);
void ora_fc2gp_dft(double *a, int n, int lot);

/* ---- one member round trip (the unit of work of the benchmark) ----
 * sp_in : nfl*2*ncsp spectral field-levels (vorticity, divergence,
 *         temperature at nlev levels each, then ln(ps)): nfl = 3*nlev+1
 * gp    : grid outputs, (nfl + 1 + 2*nlev) fields of nlat*nlon:
 *         the nfl fields, d(lnps)/dmu, then u*cos, v*cos at nlev levels
 * sp_out: nfl*2*ncsp spectral coefficients after gp2fc+fc2sp
 * work  : ora_roundtrip_worksize() doubles */
long ora_roundtrip_worksize(int ntru, int nlat, int nlon, int nlev);
int ora_roundtrip(int ntru, int nlat, int nlon, int nlev, const double *tables,
                  const double *skspgp, const double *skgpsp, double plavor,
                  int use_dft, const double *sp_in, double *gp,
                  double *sp_out, double *work);

/* ---- semi-implicit solve (src/plasim.f90:1683-1888, 2841-2876) ---- */
int ora_minvers(double *a, int n);
void ora_makebm(int nlev, int ntru, double delt, const double *t0,
                const double *dsigma, const double *g, const double *tau,
                double *bm1);
void ora_si_apply(int nlev, int nsp, const int *nindex, double delt, const double *t0,
                  const double *dsigma, const double *g, const double *tau,
                  const double *bm1, const double *sop, const double *spm,
                  const double *atm, const double *adm, double *stt, double *sdt,
                  double *spt);

/* ---- grid-point dynamics (exo_oracle_dyn.c) ---- */
void ora_initsi(int nlev, const double *sigmah, const double *dsigma, const double *rdsig,
                const double *t0, double akap, double *g, double *tau, double *c,
                double *t01s2, double *tkp);
void ora_calcgp(int nlev, long nhor, const double *gu, const double *gv, const double *gt,
                const double *gd, const double *gz, const double *gq, const double *gp,
                const double *gpj, const double *gpm, const double *dq, const double *rcsq,
                const double *sigmah, const double *dsigma, const double *rdsig,
                const double *t0, const double *t01s2, const double *tkp, const double *c,
                const double *g, const double *scal, int nqspec, double *gtn, double *gqn,
                double *guz, double *gvz, double *gvp, double *gphi, double *dw);

/* ---- long-wave radiation column (src/radmod.f90:2122-2456) ----
 * arrays [col][lev]; dt has nlev+1 values per column (surface last),
 * dq/dqo3/dqco2/dcc/dql nlev; dftu/dftd out nlev+1 */
void ora_lwr(int nlev, int ncol, const double *sigma, const double *sigmah,
             const double *dsigma, double ga, double th2oc, double tpofmt,
             double acllwr, double clgray, int npbroaden, const double *dp,
             const double *dt, const double *dq, const double *dqo3,
             const double *dqco2, const double *dcc, const double *dql,
             const double *dls, double *dftu, double *dftd);
/* heating rates in K/s (src/radmod.f90:1107-1108) */
void ora_lwr_heating(int nlev, int ncol, const double *dsigma, double ga, double acpd,
                     double adv, const double *dp, const double *dq, const double *dftu,
                     const double *dftd, double *dtdt);
/* sigma levels for neqsig = 0 (src/plasim.f90:1553-1567) */
void ora_sigma_levels(int nlev, double *sigma, double *sigmah, double *dsigma);

#ifdef __cplusplus
}
#endif
#endif
