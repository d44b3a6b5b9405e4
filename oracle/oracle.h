/*
 * oracle.h -- FP64 CPU restatement of the reference's dense tensor-decomposition
 * algorithms (891922758/first-project, src/ and decomposition/), used ONLY as the
 * correctness checker by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg.  Nothing in the product library links or calls this code.
 *
 * Conventions (reference src/tensor2matrix.cu:4-36, src/cp_als.cu:5-6):
 *   tensors are column-major, X[i + j*I + k*I*J (+ l*I*J*K)];
 *   factor matrices are column-major I x R, U[i + r*I];
 *   mode-n unfoldings order the remaining indices with the earliest mode fastest.
 *
 * Parity pinning per function is listed in oracle/README.md.
 */
#ifndef TDX_ORACLE_H
#define TDX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based generator: Philox-4x32-10 (Salmon et al., SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* value n of stream (seed, stream): U(0,1) with 23 random bits, (m + 1/2) 2^-23, never 0 or 1 */
float orc_uniform(uint64_t seed, uint32_t stream, uint64_t n);
/* N(0,1) by Box-Muller from the uniform pair (2n, 2n+1) of the stream, evaluated in FP64 */
double orc_normal(uint64_t seed, uint32_t stream, uint64_t n);
void orc_fill_uniform(float *out, long n, uint64_t seed, uint32_t stream);
void orc_fill_normal(float *out, long n, uint64_t seed, uint32_t stream);

/* ---- elementwise / permutation primitives (exact in FP32) ---- */
void orc_t2m(const float *T1, float *T2, long flag, long m, long n, long k);           /* src/tensor2matrix.cu:4-52 */
void orc_kr(const float *A, const float *B, float *AkrB, long a, long b, long r);       /* src/kr.cu:3-11 */
void orc_hadamard(const float *A, const float *B, float *C, long n);                    /* src/hadmard.cu:2-15 */
void orc_hadamard3(const float *A, const float *B, const float *C, float *D, long n);   /* src/hadmard.cu:17-30 */
void orc_tensor_product(const float *A, const float *B, float *C, long a, long b, long c, long d); /* src/tensorproduct.cu:4-14 */
void orc_tminust(const float *L, float *R, long n);                                     /* src/tminust.cu:2-15 */
void orc_identity(float *A, long m);                                                    /* src/tminust.cu:17-36 */

/* ---- storage formats ---- */
uint16_t orc_f32_to_bf16(float x);      /* IEEE round-to-nearest-even, NaN kept quiet */
float orc_bf16_to_f32(uint16_t h);
uint16_t orc_f32_to_f16(float x);       /* IEEE binary16, RNE (src/float2half.cu:4-12 semantics) */
uint8_t orc_f32_to_e4m3(float x);       /* OCP FP8 E4M3FN, RNE, saturating to +-448 */
float orc_e4m3_to_f32(uint8_t q);
/* per-mode-1-slice scale: s_i = max_{j,k}|X(i,:,..)| / 448 (1 if the slice is all zero),
 * q = e4m3(x / s_i) */
void orc_quantize_e4m3_mode1(const float *X, long I, long rest, uint8_t *Q, float *scales);

/* ---- FP64 multilinear algebra ---- */
/* MTTKRP of an order-d tensor (d<=4) on mode n (0-based):
 * out[i_n + r*In] = sum_{other idx} X[...] * prod_{m!=n} U_m[i_m + r*I_m]
 * X given as float (exact data), accumulated in double; U given in double. */
void orc_mttkrp(const float *X, int order, const long *dims, const double *const *U,
                long R, int mode, double *out);
/* same with per-mode-1-slice scales applied to X (X_eff[i,...] = s_i * Xq) and Xq given
 * as float values (dequantised codes) */
void orc_mttkrp_scaled(const float *Xq, const float *scales, int order, const long *dims,
                       const double *const *U, long R, int mode, double *out);
/* reference ttm(X,U,XU,flag,a,b,c,d): X (a x b x c) times U (d x mode-size, col-major)
 * along mode flag (1..3); output d x b x c, a x d x c or a x b x d (src/ttm.cu:4-47) */
void orc_ttm(const double *X, const double *U, double *XU, long flag, long a, long b, long c, long d);
/* general order-d TTM with transposed factor: Y = X x_n U^T with U (I_n x r) col-major */
void orc_ttm_t(const double *X, int order, const long *dims, int mode, const double *U, long r, double *Y);
void orc_gram(const double *U, long I, long R, double *G);                  /* G = U^T U */
/* Cholesky of SPD V (R x R col-major) in place (lower); returns 0 or the failing column+1 */
int orc_cholesky(double *V, long R);
/* solve U (I x R) := M V^{-1} with V = L L^T (lower factor from orc_cholesky) */
void orc_chol_solve_rows(const double *L, long R, const double *M, long I, double *U);
/* symmetric pseudo-inverse solve via Jacobi eigendecomposition (fallback for singular V) */
void orc_pinv_solve_rows(const double *V, long R, const double *M, long I, double *U, double rcond);
/* cyclic Jacobi eigensolver: A (n x n col-major, symmetric) -> ascending eigenvalues w and
 * eigenvectors V (columns), sign fixed so the largest-|.| entry of each vector is positive */
void orc_sym_eig(const double *A, long n, double *w, double *V);

/* ---- CP ---- */
/* X = sum_r prod_m U_m[:,r] (reference gencp/gencp4, src/rse.cu:17-30,53-73), double out */
void orc_cp_reconstruct(int order, const long *dims, const double *const *U, long R, double *Xhat);
/* relative error ||X - Xhat|| / ||X|| by a direct compensated residual pass */
double orc_cp_relerr(const float *X, int order, const long *dims, const double *const *U, long R);
/* ALS: U[m] hold the init on entry and the result on exit (the mode-0 init is not read,
 * as in src/cp_als.cu:44-63 where A is first written).  Records the per-sweep fast-fit
 * relative error (Tensorlab cpd_als.m:281-283 identity) and, if relerr_direct != NULL,
 * the direct residual relative error after every sweep.  Returns 0 or a solve status. */
int orc_cp_als(const float *X, int order, const long *dims, long R, double **U, int sweeps,
               double *relerr_fit, double *relerr_direct);

/* ---- Tucker ---- */
/* HOSVD (src/tucker_hosvd.cu:3-81, src/tucker4_hosvd.cu:3-121): U_n = eigenvectors of the
 * r_n largest eigenvalues of X(n) X(n)^T, columns in ascending-eigenvalue order (the
 * reference copies the last r columns of the ascending syevdx output); core = X x_n U_n^T. */
void orc_tucker_hosvd(const float *X, int order, const long *dims, const long *ranks,
                      double **U, double *core);
/* HOOI (decomposition/GPU/normal/tucker/tucker_hooi.cu:103-168 generalised to order d):
 * U holds the HOSVD init on entry. */
void orc_tucker_hooi(const float *X, int order, const long *dims, const long *ranks,
                     int iters, double **U, double *core, double *relerr_hist);
/* Tucker reconstruction core x_1 U_1 ... x_d U_d (src/rse.cu:32-51,75-96) */
void orc_tucker_reconstruct(int order, const long *dims, const long *ranks,
                            const double *core, const double *const *U, double *Xhat);

/* ---- MLP (3 layers, GELU) for the learned ALS extrapolation ---- */
/* y = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3 ; W row-major (out x in) */
void orc_mlp_forward(const double *x, long batch, long nin, long width, long nout,
                     const double *W1, const double *b1, const double *W2, const double *b2,
                     const double *W3, const double *b3, double *y);
/* loss = mean((y - t)^2)/2 over batch*nout; gradients accumulated into g* (overwritten) */
double orc_mlp_backward(const double *x, const double *t, long batch, long nin, long width, long nout,
                        const double *W1, const double *b1, const double *W2, const double *b2,
                        const double *W3, const double *b3,
                        double *gW1, double *gb1, double *gW2, double *gb2, double *gW3, double *gb3);
/* Adam (Kingma & Ba 2015, Alg. 1) in place; step t is 1-based */
void orc_adam(double *p, const double *g, double *m, double *v, long n, long t,
              double lr, double b1, double b2, double eps);

#ifdef __cplusplus
}
#endif
#endif
