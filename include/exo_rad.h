/*
 * exo_rad.h -- C ABI of the long-wave radiation column and of the learned
 * emulator's kernels (part of libexosht.so).
 *
 * exo_lwr restates the reference's long-wave column (subroutine lwr,
 * exoplasim/plasim/src/radmod.f90:2122-2456) with its heating rate
 * (src/radmod.f90:1100-1108) on the GPU, one thread per column.  Column
 * arrays are device pointers in the reference layout field(NHOR, lev): the
 * column index is fastest.  fp64 selects double arrays, else float.
 *
 * The exo_mlp_* functions are the forward and backward kernels of the
 * emulator network (BF16 operands, FP32 accumulation on MFMA); they have no
 * counterpart in the reference, which has no emulator.
 */
#ifndef EXO_RAD_H
#define EXO_RAD_H
#ifdef __cplusplus
extern "C" {
#endif

#define EXO_RAD_OK 0
#define EXO_RAD_ERR_ARG (-1)
#define EXO_RAD_ERR_HIP (-5)

typedef struct {
  int nlev;                       /* 2..64 */
  const double *sigma, *sigmah, *dsigma; /* device, nlev each (pumamod) */
  double ga;                      /* gravity (src/p_earth.f90:58) */
  double th2oc, tpofmt, acllwr, clgray; /* radmod_nl (src/radmod.f90:37-43) */
  int npbroaden;                  /* src/radmod.f90:71 */
  double acpd, adv;               /* gascon/akap, ACPV/acpd - 1 */
} exo_lwr_params;

typedef struct {
  const void *dp;                 /* (NHOR)        surface pressure, Pa */
  const void *dt;                 /* (NHOR,NLEP)   temperature, surface last */
  const void *dq;                 /* (NHOR,NLEV)   specific humidity */
  const void *dqo3, *dqco2;       /* (NHOR,NLEV)   ozone kg/kg, co2 ppmv */
  const void *dcc, *dql;          /* (NHOR,NLEV)   cloud cover, liquid water */
  const void *dls;                /* (NHOR)        land/sea mask */
} exo_lwr_columns;

typedef struct {
  void *dftu, *dftd;              /* (NHOR,NLEP) upward / downward flux or NULL */
  void *heat;                     /* (NHOR,NLEV) long-wave heating, K/day, or NULL */
  void *olr;                      /* (NHOR) outgoing long-wave radiation or NULL */
} exo_lwr_out;

int exo_lwr(const exo_lwr_params *p, int fp64, int ncol, const exo_lwr_columns *in,
            exo_lwr_out *out, void *stream);

/* emulator layers: X[M][K], W[N][K], Wt[K][N] bf16 (uint16 storage),
 * K, N multiples of 32 (N of a hidden layer a multiple of 64); nvalid/kvalid
 * are the unpadded sizes.  act != 0: Z (fp32, may be NULL) and H = gelu(Z)
 * (bf16); act == 0: Y = Z (fp32). */
int exo_mlp_forward(const void *X, const void *W, const float *b, int M, int K, int N,
                    int nvalid, int act, float *Z, void *H, float *Y, void *stream);
/* dZprev = (dZ W) * gelu'(Zprev): bf16 (and fp32 if dZprev32 != NULL) */
int exo_mlp_backward_dx(const void *dZ, const void *Wt, const float *Zprev, int M, int N, int K,
                        int kvalid, void *dZprev, float *dZprev32, void *stream);
/* dW = dZ^T X (split over nsplit batch chunks into part[nsplit][N][K]) and
 * db = sum over the batch of dZ32 */
int exo_mlp_backward_dw(const void *dZ, const void *X, const float *dZ32, int M, int N, int K,
                        int nsplit, float *part, float *dW, float *db, void *stream);
/* FP32 master weights w[nout][nin] -> zero-padded bf16 W[N][K], Wt[K][N] */
int exo_mlp_cast_weights(const float *w, int nout, int nin, int N, int K, void *Wb, void *Wtb,
                         void *stream);
/* fp32 rows src[M][nin] -> zero-padded bf16 dst[M][K] */
int exo_mlp_cast_rows(const float *src, int M, int nin, int K, void *dst, void *stream);

/* A trained emulator for inference: bf16 weights W1[256][64], W2[256][256],
 * W3[256][256], W4[32][256] (zero padded, as exo_mlp_cast_weights writes
 * them), fp32 biases, and the output normalisation ymean/ystd (21 values).
 * nlev must be 20. */
typedef struct {
  int nlev;
  const void *w[4];
  const float *b[4];
  const float *ymean, *ystd;
} exo_lw_net;
/* The learned long-wave column in one kernel, in place of lwr
 * (src/radmod.f90:2122) and the heating rate (src/radmod.f90:1107-1108):
 * from fp32 device columns in the model layout field(NHOR, lev) -- dt
 * (nlev+1 rows, the last the surface), dq (nlev rows), dp (1 row), dqco2 (its
 * first row) -- to heat (nlev rows, K/day) and olr (1 row, W/m2).  Features,
 * the four layers (bf16 MFMA, fp32 accumulate) and the output scaling run
 * per block of 32 columns with the activations in LDS. */
int exo_lw_emulate(const exo_lw_net *net, int ncol, const float *dt, const float *dq,
                   const float *dp, const float *dqco2, float *heat, float *olr, void *stream);
/* The same, and outside[col] = 1 for every column with one of its 43
 * network inputs outside [flo, fhi] (the training envelope) or not a
 * number, else 0: the columns the caller hands to the plain column exo_lwr
 * instead, as the reference would have computed all of them. */
int exo_lw_emulate_guarded(const exo_lw_net *net, const float *flo, const float *fhi, int ncol,
                           const float *dt, const float *dq, const float *dp,
                           const float *dqco2, float *heat, float *olr, int *outside,
                           void *stream);

#ifdef __cplusplus
}
#endif
#endif
