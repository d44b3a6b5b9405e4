/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.  FP64 CPU restatement of the reference's
 * algorithms; the checker for the HIP library, never linked into it.
 * Each function cites the reference file:line it restates.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11).  Multipliers and Weyl key increments from the paper. */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

static inline uint32_t philox_word(uint64_t seed, uint32_t stream, uint64_t n) {
    uint64_t blk = n >> 2;
    uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), stream, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return o[n & 3];
}

float orc_uniform(uint64_t seed, uint32_t stream, uint64_t n) {
    uint32_t w = philox_word(seed, stream, n);
    /* 23 random bits, centred in their bucket: (m + 0.5) * 2^-23, exact in float, in (0,1) */
    return ((float)(w >> 9) + 0.5f) * (1.0f / 8388608.0f);
}

double orc_normal(uint64_t seed, uint32_t stream, uint64_t n) {
    double u1 = (double)orc_uniform(seed, stream, 2 * n);
    double u2 = (double)orc_uniform(seed, stream, 2 * n + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void orc_fill_uniform(float *out, long n, uint64_t seed, uint32_t stream) {
    for (long i = 0; i < n; ++i) out[i] = orc_uniform(seed, stream, (uint64_t)i);
}

void orc_fill_normal(float *out, long n, uint64_t seed, uint32_t stream) {
    for (long i = 0; i < n; ++i) out[i] = (float)orc_normal(seed, stream, (uint64_t)i);
}

/* ------------------------------------------------------------------------- */
/* t2m: src/tensor2matrix.cu:4-36.  Element (row,col,tube) of the m x n x k tensor
 * goes to  mode 1: tube*m*n + col*m + row   (identity)
 *          mode 2: tube*m*n + row*n + col   (X(2) = n x (m k), column row + tube*m)
 *          mode 3: k*(col*m+row) + tube     (X(3) = k x (m n), column row + col*m)   */
void orc_t2m(const float *T1, float *T2, long flag, long m, long n, long k) {
    for (long tube = 0; tube < k; ++tube)
        for (long col = 0; col < n; ++col)
            for (long row = 0; row < m; ++row) {
                long src = tube * m * n + col * m + row;
                long dst;
                if (flag == 1) dst = src;
                else if (flag == 2) dst = tube * m * n + row * n + col;
                else dst = k * (col * m + row) + tube;
                T2[dst] = T1[src];
            }
}

/* kr: src/kr.cu:9 -- column t of A (.) B is kron(A(:,t), B(:,t)), B index fastest:
 * AkrB[jb + ia*b + t*a*b] = A[ia + t*a] * B[jb + t*b] (one FP32 product) */
void orc_kr(const float *A, const float *B, float *AkrB, long a, long b, long r) {
    for (long t = 0; t < r; ++t)
        for (long ia = 0; ia < a; ++ia)
            for (long jb = 0; jb < b; ++jb)
                AkrB[jb + ia * b + t * a * b] = A[ia + t * a] * B[jb + t * b];
}

void orc_hadamard(const float *A, const float *B, float *C, long n) {
    for (long i = 0; i < n; ++i) C[i] = A[i] * B[i];
}

/* src/hadmard.cu:20: res = M*N*K evaluated left to right */
void orc_hadamard3(const float *A, const float *B, const float *C, float *D, long n) {
    for (long i = 0; i < n; ++i) {
        float t = A[i] * B[i];
        D[i] = t * C[i];
    }
}

/* tensor_product: src/tensorproduct.cu:9-11 -- Kronecker A (a x b) (x) B (c x d),
 * C ((a c) x (b d)) col-major: C[(ia*c + ic) + (jb*d + jd)*a*c] = A[ia + jb*a] * B[ic + jd*c] */
void orc_tensor_product(const float *A, const float *B, float *C, long a, long b, long c, long d) {
    for (long jb = 0; jb < b; ++jb)
        for (long jd = 0; jd < d; ++jd)
            for (long ia = 0; ia < a; ++ia)
                for (long ic = 0; ic < c; ++ic)
                    C[(ia * c + ic) + (jb * d + jd) * a * c] = A[ia + jb * a] * B[ic + jd * c];
}

/* src/tminust.cu:5: right := left - right */
void orc_tminust(const float *L, float *R, long n) {
    for (long i = 0; i < n; ++i) R[i] = L[i] - R[i];
}

void orc_identity(float *A, long m) {
    for (long c = 0; c < m; ++c)
        for (long r = 0; r < m; ++r) A[c * m + r] = (r == c) ? 1.0f : 0.0f;
}

/* ------------------------------------------------------------------------- */
static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

uint16_t orc_f32_to_bf16(float x) {
    uint32_t u = f2u(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u); /* quiet NaN */
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

float orc_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

/* IEEE binary16, round to nearest even (the __float2half rounding of src/float2half.cu:8) */
uint16_t orc_f32_to_f16(float x) {
    uint32_t u = f2u(x);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t ax = u & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to inf */
    if (ax < 0x38800000u) { /* result subnormal or zero: quantum 2^-24 */
        double v = (double)u2f(ax) * 16777216.0; /* exact */
        double fl = floor(v);
        double rem = v - fl;
        uint32_t m = (uint32_t)fl;
        if (rem > 0.5 || (rem == 0.5 && (m & 1u))) m++;
        return (uint16_t)(sign | m);
    }
    uint32_t e = (ax >> 23) - 127 + 15;
    uint32_t mant = ax & 0x7fffffu;
    uint32_t h = (e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

/* OCP 8-bit floating point spec v1.0, E4M3FN: bias 7, no infinities, max 448,
 * subnormals m * 2^-9.  Round to nearest even, saturate finite overflow to 448, NaN->0x7f. */
uint8_t orc_f32_to_e4m3(float x) {
    uint32_t u = f2u(x);
    uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    float ax = fabsf(x);
    if (isnan(x)) return (uint8_t)(sign | 0x7f);
    if (ax >= 464.0f) return (uint8_t)(sign | 0x7e); /* 448 = 1.75*2^8; >=464 would round past it */
    if (ax < 0.015625f) { /* below min normal 2^-6: quantum 2^-9 */
        double v = (double)ax * 512.0;
        double fl = floor(v);
        double rem = v - fl;
        uint32_t m = (uint32_t)fl;
        if (rem > 0.5 || (rem == 0.5 && (m & 1u))) m++;
        return (uint8_t)(sign | m); /* m == 8 gives 0x08 = 2^-6, the min normal: correct */
    }
    uint32_t au = f2u(ax);
    int e = (int)((au >> 23) & 0xff) - 127; /* -6..8 */
    uint32_t mant = au & 0x7fffffu;
    uint32_t q = ((uint32_t)(e + 7) << 3) | (mant >> 20);
    uint32_t rem = mant & 0xfffffu;
    if (rem > 0x80000u || (rem == 0x80000u && (q & 1u))) q++;
    if (q > 0x7e) q = 0x7e;
    return (uint8_t)(sign | q);
}

float orc_e4m3_to_f32(uint8_t q) {
    float s = (q & 0x80) ? -1.0f : 1.0f;
    uint32_t e = (q >> 3) & 0xf, m = q & 7;
    if (e == 0xf && m == 7) return NAN;
    if (e == 0) return s * (float)m * (1.0f / 512.0f);
    return s * ldexpf(1.0f + (float)m / 8.0f, (int)e - 7);
}

void orc_quantize_e4m3_mode1(const float *X, long I, long rest, uint8_t *Q, float *scales) {
    for (long i = 0; i < I; ++i) {
        float amax = 0.0f;
        for (long t = 0; t < rest; ++t) {
            float a = fabsf(X[i + t * I]);
            if (a > amax) amax = a;
        }
        scales[i] = (amax > 0.0f) ? amax / 448.0f : 1.0f;
    }
    for (long t = 0; t < rest; ++t)
        for (long i = 0; i < I; ++i) Q[i + t * I] = orc_f32_to_e4m3(X[i + t * I] / scales[i]);
}

/* ------------------------------------------------------------------------- */
/* MTTKRP (src/cp_als.cu:46-52,66-69,85-88 and src/mttkrp.cu:11-77 for order 3;
 * src/cp4_als.cu:67-73,91-96,112-117,135-138 for order 4): the mode-n unfolding times
 * the Khatri-Rao product of the other factors, earliest mode fastest.
 * Computed fibre by fibre along mode 0 with FP64 accumulation. */
static void mttkrp_core(const float *X, const float *scales, int order, const long *dims,
                        const double *const *U, long R, int mode, double *out) {
    long I0 = dims[0];
    long nfib = 1;
    for (int m = 1; m < order; ++m) nfib *= dims[m];
    long In = dims[mode];
    memset(out, 0, sizeof(double) * In * R);
#pragma omp parallel
    {
        double *acc = (double *)calloc((size_t)(In * R), sizeof(double));
        double *w = (double *)malloc(sizeof(double) * R);
#pragma omp for schedule(static)
        for (long f = 0; f < nfib; ++f) {
            long idx[4] = {0, 0, 0, 0};
            long rem = f;
            for (int m = 1; m < order; ++m) { idx[m] = rem % dims[m]; rem /= dims[m]; }
            const float *x = X + f * I0;
            for (long r = 0; r < R; ++r) {
                double p = 1.0;
                for (int m = 1; m < order; ++m)
                    if (m != mode) p *= U[m][idx[m] + r * dims[m]];
                w[r] = p;
            }
            if (mode == 0) {
                for (long r = 0; r < R; ++r) {
                    double wr = w[r];
                    double *a = acc + r * In;
                    if (scales)
                        for (long i = 0; i < I0; ++i) a[i] += (double)scales[i] * (double)x[i] * wr;
                    else
                        for (long i = 0; i < I0; ++i) a[i] += (double)x[i] * wr;
                }
            } else {
                long in = idx[mode];
                const double *U0 = U[0];
                for (long r = 0; r < R; ++r) {
                    double t = 0.0;
                    const double *u = U0 + r * I0;
                    if (scales)
                        for (long i = 0; i < I0; ++i) t += (double)scales[i] * (double)x[i] * u[i];
                    else
                        for (long i = 0; i < I0; ++i) t += (double)x[i] * u[i];
                    acc[in + r * In] += t * w[r];
                }
            }
        }
#pragma omp critical
        for (long e = 0; e < In * R; ++e) out[e] += acc[e];
        free(acc);
        free(w);
    }
}

void orc_mttkrp(const float *X, int order, const long *dims, const double *const *U,
                long R, int mode, double *out) {
    mttkrp_core(X, NULL, order, dims, U, R, mode, out);
}

void orc_mttkrp_scaled(const float *Xq, const float *scales, int order, const long *dims,
                       const double *const *U, long R, int mode, double *out) {
    mttkrp_core(Xq, scales, order, dims, U, R, mode, out);
}

/* ttm: src/ttm.cu:21,29,37.  U is d x (mode size), col-major.
 *  flag 1: XU (d x b x c)[p + j*d + k*d*b] = sum_i U[p + i*d] X[i,j,k]
 *  flag 2: XU (a x d x c)[i + p*a + k*a*d] = sum_j U[p + j*d] X[i,j,k]
 *  flag 3: XU (a x b x d)[i + j*a + p*a*b] = sum_k U[p + k*d] X[i,j,k]              */
void orc_ttm(const double *X, const double *U, double *XU, long flag, long a, long b, long c, long d) {
    long osz = (flag == 1 ? d * b * c : flag == 2 ? a * d * c : a * b * d);
    memset(XU, 0, sizeof(double) * osz);
    for (long k = 0; k < c; ++k)
        for (long j = 0; j < b; ++j)
            for (long i = 0; i < a; ++i) {
                double x = X[i + j * a + k * a * b];
                for (long p = 0; p < d; ++p) {
                    if (flag == 1) XU[p + j * d + k * d * b] += U[p + i * d] * x;
                    else if (flag == 2) XU[i + p * a + k * a * d] += U[p + j * d] * x;
                    else XU[i + j * a + p * a * b] += U[p + k * d] * x;
                }
            }
}

/* Y = X x_mode U^T, U (I_mode x r) col-major; Y has dims with dims[mode] replaced by r */
void orc_ttm_t(const double *X, int order, const long *dims, int mode, const double *U, long r, double *Y) {
    long P = 1, Q = 1;
    for (int m = 0; m < mode; ++m) P *= dims[m];
    for (int m = mode + 1; m < order; ++m) Q *= dims[m];
    long In = dims[mode];
#pragma omp parallel for schedule(static)
    for (long q = 0; q < Q; ++q)
        for (long c = 0; c < r; ++c) {
            double *y = Y + q * P * r + c * P;
            for (long p = 0; p < P; ++p) y[p] = 0.0;
            for (long i = 0; i < In; ++i) {
                double u = U[i + c * In];
                const double *x = X + q * P * In + i * P;
                for (long p = 0; p < P; ++p) y[p] += u * x[p];
            }
        }
}

void orc_gram(const double *U, long I, long R, double *G) {
    for (long a = 0; a < R; ++a)
        for (long b = 0; b < R; ++b) {
            double s = 0.0;
            for (long i = 0; i < I; ++i) s += U[i + a * I] * U[i + b * I];
            G[a + b * R] = s;
        }
}

/* Cholesky-Banachiewicz (Golub & Van Loan, Alg. 4.2.1), lower triangle in place */
int orc_cholesky(double *V, long R) {
    for (long j = 0; j < R; ++j) {
        double d = V[j + j * R];
        for (long k = 0; k < j; ++k) d -= V[j + k * R] * V[j + k * R];
        if (!(d > 0.0)) return (int)(j + 1);
        d = sqrt(d);
        V[j + j * R] = d;
        for (long i = j + 1; i < R; ++i) {
            double s = V[i + j * R];
            for (long k = 0; k < j; ++k) s -= V[i + k * R] * V[j + k * R];
            V[i + j * R] = s / d;
        }
        for (long i = 0; i < j; ++i) V[i + j * R] = 0.0;
    }
    return 0;
}

/* The normal equations of src/cp_als.cu:52-63: U V = M with V symmetric; row by row,
 * forward then backward substitution with L L^T = V. */
void orc_chol_solve_rows(const double *L, long R, const double *M, long I, double *U) {
    double *y = (double *)malloc(sizeof(double) * R);
    for (long i = 0; i < I; ++i) {
        for (long a = 0; a < R; ++a) {
            double s = M[i + a * I];
            for (long k = 0; k < a; ++k) s -= L[a + k * R] * y[k];
            y[a] = s / L[a + a * R];
        }
        for (long a = R - 1; a >= 0; --a) {
            double s = y[a];
            for (long k = a + 1; k < R; ++k) s -= L[k + a * R] * y[k];
            y[a] = s / L[a + a * R];
        }
        for (long a = 0; a < R; ++a) U[i + a * I] = y[a];
    }
    free(y);
}

/* cyclic Jacobi (Golub & Van Loan, Alg. 8.5.3) */
void orc_sym_eig(const double *Ain, long n, double *w, double *V) {
    double *A = (double *)malloc(sizeof(double) * n * n);
    memcpy(A, Ain, sizeof(double) * n * n);
    for (long i = 0; i < n * n; ++i) V[i] = 0.0;
    for (long i = 0; i < n; ++i) V[i + i * n] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (long q = 0; q < n; ++q)
            for (long p = 0; p < n; ++p) {
                double v = A[p + q * n] * A[p + q * n];
                tot += v;
                if (p != q) off += v;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (long p = 0; p < n - 1; ++p)
            for (long q = p + 1; q < n; ++q) {
                double apq = A[p + q * n];
                if (apq == 0.0) continue;
                double app = A[p + p * n], aqq = A[q + q * n];
                double tau = (aqq - app) / (2.0 * apq);
                double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                for (long k = 0; k < n; ++k) { /* A := A J (columns p,q) */
                    double akp = A[k + p * n], akq = A[k + q * n];
                    A[k + p * n] = c * akp - s * akq;
                    A[k + q * n] = s * akp + c * akq;
                }
                for (long k = 0; k < n; ++k) { /* A := J^T A (rows p,q) */
                    double apk = A[p + k * n], aqk = A[q + k * n];
                    A[p + k * n] = c * apk - s * aqk;
                    A[q + k * n] = s * apk + c * aqk;
                }
                for (long k = 0; k < n; ++k) {
                    double vkp = V[k + p * n], vkq = V[k + q * n];
                    V[k + p * n] = c * vkp - s * vkq;
                    V[k + q * n] = s * vkp + c * vkq;
                }
            }
    }
    /* sort ascending (selection), fix signs */
    for (long i = 0; i < n; ++i) w[i] = A[i + i * n];
    for (long i = 0; i < n; ++i) {
        long mi = i;
        for (long j = i + 1; j < n; ++j) if (w[j] < w[mi]) mi = j;
        if (mi != i) {
            double tw = w[i]; w[i] = w[mi]; w[mi] = tw;
            for (long k = 0; k < n; ++k) { double t = V[k + i * n]; V[k + i * n] = V[k + mi * n]; V[k + mi * n] = t; }
        }
    }
    for (long i = 0; i < n; ++i) {
        long am = 0;
        for (long k = 1; k < n; ++k) if (fabs(V[k + i * n]) > fabs(V[am + i * n])) am = k;
        if (V[am + i * n] < 0) for (long k = 0; k < n; ++k) V[k + i * n] = -V[k + i * n];
    }
    free(A);
}

void orc_pinv_solve_rows(const double *Vm, long R, const double *M, long I, double *U, double rcond) {
    double *w = (double *)malloc(sizeof(double) * R);
    double *Q = (double *)malloc(sizeof(double) * R * R);
    double *P = (double *)calloc((size_t)(R * R), sizeof(double));
    orc_sym_eig(Vm, R, w, Q);
    double wmax = 0.0;
    for (long a = 0; a < R; ++a) if (fabs(w[a]) > wmax) wmax = fabs(w[a]);
    for (long e = 0; e < R; ++e) {
        if (fabs(w[e]) <= rcond * wmax) continue;
        double inv = 1.0 / w[e];
        for (long b = 0; b < R; ++b)
            for (long a = 0; a < R; ++a) P[a + b * R] += Q[a + e * R] * Q[b + e * R] * inv;
    }
    for (long i = 0; i < I; ++i)
        for (long b = 0; b < R; ++b) {
            double s = 0.0;
            for (long a = 0; a < R; ++a) s += M[i + a * I] * P[a + b * R];
            U[i + b * I] = s;
        }
    free(w); free(Q); free(P);
}

/* ------------------------------------------------------------------------- */
void orc_cp_reconstruct(int order, const long *dims, const double *const *U, long R, double *Xhat) {
    long N = 1;
    for (int m = 0; m < order; ++m) N *= dims[m];
#pragma omp parallel for schedule(static)
    for (long e = 0; e < N; ++e) {
        long idx[4] = {0, 0, 0, 0}, rem = e;
        for (int m = 0; m < order; ++m) { idx[m] = rem % dims[m]; rem /= dims[m]; }
        double s = 0.0;
        for (long r = 0; r < R; ++r) {
            double p = 1.0;
            for (int m = 0; m < order; ++m) p *= U[m][idx[m] + r * dims[m]];
            s += p;
        }
        Xhat[e] = s;
    }
}

/* src/rse.cu:4-15 computes ||X - X1|| / ||X|| with two FP32 nrm2 calls; here the
 * residual and the norm are accumulated in FP64 with Kahan compensation per thread. */
double orc_cp_relerr(const float *X, int order, const long *dims, const double *const *U, long R) {
    long N = 1;
    for (int m = 0; m < order; ++m) N *= dims[m];
    double res = 0.0, nx = 0.0;
#pragma omp parallel
    {
        double rs = 0.0, rc = 0.0, xs = 0.0, xc = 0.0;
#pragma omp for schedule(static)
        for (long e = 0; e < N; ++e) {
            long idx[4] = {0, 0, 0, 0}, rem = e;
            for (int m = 0; m < order; ++m) { idx[m] = rem % dims[m]; rem /= dims[m]; }
            double s = 0.0;
            for (long r = 0; r < R; ++r) {
                double p = 1.0;
                for (int m = 0; m < order; ++m) p *= U[m][idx[m] + r * dims[m]];
                s += p;
            }
            double d = (double)X[e] - s;
            double y = d * d - rc, t = rs + y;
            rc = (t - rs) - y; rs = t;
            double x2 = (double)X[e] * (double)X[e];
            y = x2 - xc; t = xs + y; xc = (t - xs) - y; xs = t;
        }
#pragma omp critical
        { res += rs; nx += xs; }
    }
    return sqrt(res / nx);
}

/* CP-ALS, src/cp_als.cu:44-119 (order 3) and src/cp4_als.cu:64-173 (order 4):
 * for each mode n in order: M = MTTKRP_n, V = hadamard of the other Grams, U_n = M V^{-1}.
 * The reference solves with LU (getrf/getrs); V is SPD so Cholesky gives the same
 * solution; a pseudo-inverse replaces it when Cholesky fails. */
int orc_cp_als(const float *X, int order, const long *dims, long R, double **U, int sweeps,
               double *relerr_fit, double *relerr_direct) {
    long N = 1, Imax = 0;
    for (int m = 0; m < order; ++m) { N *= dims[m]; if (dims[m] > Imax) Imax = dims[m]; }
    double normx2 = 0.0;
    for (long e = 0; e < N; ++e) normx2 += (double)X[e] * (double)X[e];
    double *G[4];
    for (int m = 0; m < order; ++m) {
        G[m] = (double *)malloc(sizeof(double) * R * R);
        orc_gram(U[m], dims[m], R, G[m]);
    }
    double *M = (double *)malloc(sizeof(double) * Imax * R);
    double *V = (double *)malloc(sizeof(double) * R * R);
    int status = 0;
    for (int it = 0; it < sweeps; ++it) {
        for (int n = 0; n < order; ++n) {
            orc_mttkrp(X, order, dims, (const double *const *)U, R, n, M);
            for (long e = 0; e < R * R; ++e) {
                double v = 1.0;
                for (int m = 0; m < order; ++m) if (m != n) v *= G[m][e];
                V[e] = v;
            }
            double *L = (double *)malloc(sizeof(double) * R * R);
            memcpy(L, V, sizeof(double) * R * R);
            if (orc_cholesky(L, R) == 0) {
                orc_chol_solve_rows(L, R, M, dims[n], U[n]);
            } else {
                orc_pinv_solve_rows(V, R, M, dims[n], U[n], 1e-12);
                status = 4;
            }
            free(L);
            orc_gram(U[n], dims[n], R, G[n]);
        }
        if (relerr_fit) {
            /* Tensorlab cpd_als.m:281-283: ||X - Xhat||^2 = ||X||^2 - 2<M_last,U_last> + sum(hadamard of Grams) */
            int n = order - 1;
            double ip = 0.0, hx = 0.0;
            for (long e = 0; e < dims[n] * R; ++e) ip += M[e] * U[n][e];
            for (long e = 0; e < R * R; ++e) {
                double v = 1.0;
                for (int m = 0; m < order; ++m) v *= G[m][e];
                hx += v;
            }
            double r2 = normx2 - 2.0 * ip + hx;
            relerr_fit[it] = sqrt(fabs(r2) / normx2);
        }
        if (relerr_direct) relerr_direct[it] = orc_cp_relerr(X, order, dims, (const double *const *)U, R);
    }
    for (int m = 0; m < order; ++m) free(G[m]);
    free(M); free(V);
    return status;
}

/* ------------------------------------------------------------------------- */
static void unfold_gram(const double *X, int order, const long *dims, int mode, double *G) {
    long P = 1, Q = 1;
    for (int m = 0; m < mode; ++m) P *= dims[m];
    for (int m = mode + 1; m < order; ++m) Q *= dims[m];
    long In = dims[mode];
#pragma omp parallel for schedule(dynamic)
    for (long b = 0; b < In; ++b)
        for (long a = 0; a <= b; ++a) {
            double s = 0.0;
            for (long q = 0; q < Q; ++q) {
                const double *xa = X + q * P * In + a * P;
                const double *xb = X + q * P * In + b * P;
                for (long p = 0; p < P; ++p) s += xa[p] * xb[p];
            }
            G[a + b * In] = s;
            G[b + a * In] = s;
        }
}

static void top_eigvecs(const double *G, long n, long r, double *U) {
    double *w = (double *)malloc(sizeof(double) * n);
    double *V = (double *)malloc(sizeof(double) * n * n);
    orc_sym_eig(G, n, w, V);
    /* the last r columns of the ascending decomposition (src/tucker_hosvd.cu:68-70) */
    memcpy(U, V + (n - r) * n, sizeof(double) * n * r);
    free(w); free(V);
}

static long prod(const long *d, int order) { long p = 1; for (int m = 0; m < order; ++m) p *= d[m]; return p; }

static void core_from(const double *Xd, int order, const long *dims, const long *ranks,
                      const double *const *U, double *core) {
    long cur[4];
    for (int m = 0; m < order; ++m) cur[m] = dims[m];
    const double *src = Xd;
    double *buf = NULL;
    for (int m = 0; m < order; ++m) {
        long nd[4];
        memcpy(nd, cur, sizeof(nd));
        nd[m] = ranks[m];
        double *dst = (m == order - 1) ? core : (double *)malloc(sizeof(double) * prod(nd, order));
        orc_ttm_t(src, order, cur, m, U[m], ranks[m], dst);
        if (buf) free(buf);
        buf = (m == order - 1) ? NULL : dst;
        src = dst;
        memcpy(cur, nd, sizeof(cur));
    }
}

void orc_tucker_hosvd(const float *X, int order, const long *dims, const long *ranks,
                      double **U, double *core) {
    long N = prod(dims, order);
    double *Xd = (double *)malloc(sizeof(double) * N);
    for (long e = 0; e < N; ++e) Xd[e] = X[e];
    for (int m = 0; m < order; ++m) {
        double *G = (double *)malloc(sizeof(double) * dims[m] * dims[m]);
        unfold_gram(Xd, order, dims, m, G);
        top_eigvecs(G, dims[m], ranks[m], U[m]);
        free(G);
    }
    core_from(Xd, order, dims, ranks, (const double *const *)U, core);
    free(Xd);
}

void orc_tucker_hooi(const float *X, int order, const long *dims, const long *ranks,
                     int iters, double **U, double *core, double *relerr_hist) {
    long N = prod(dims, order);
    double *Xd = (double *)malloc(sizeof(double) * N);
    double nx2 = 0.0;
    for (long e = 0; e < N; ++e) { Xd[e] = X[e]; nx2 += Xd[e] * Xd[e]; }
    for (int it = 0; it < iters; ++it) {
        for (int n = 0; n < order; ++n) {
            /* Y = X x_{m != n} U_m^T, then U_n = top eigvecs of Y(n) Y(n)^T */
            long cur[4];
            for (int m = 0; m < order; ++m) cur[m] = dims[m];
            const double *src = Xd;
            double *owned = NULL;
            for (int m = 0; m < order; ++m) {
                if (m == n) continue;
                long nd[4];
                memcpy(nd, cur, sizeof(nd));
                nd[m] = ranks[m];
                double *dst = (double *)malloc(sizeof(double) * prod(nd, order));
                orc_ttm_t(src, order, cur, m, U[m], ranks[m], dst);
                if (owned) free(owned);
                owned = dst;
                src = dst;
                memcpy(cur, nd, sizeof(cur));
            }
            double *G = (double *)malloc(sizeof(double) * dims[n] * dims[n]);
            unfold_gram(src, order, cur, n, G);
            top_eigvecs(G, dims[n], ranks[n], U[n]);
            free(G);
            if (owned) free(owned);
        }
        if (relerr_hist) {
            long nc = prod(ranks, order);
            double *c = (double *)malloc(sizeof(double) * nc);
            core_from(Xd, order, dims, ranks, (const double *const *)U, c);
            double g2 = 0.0;
            for (long e = 0; e < nc; ++e) g2 += c[e] * c[e];
            relerr_hist[it] = sqrt(fmax(nx2 - g2, 0.0) / nx2);
            free(c);
        }
    }
    core_from(Xd, order, dims, ranks, (const double *const *)U, core);
    free(Xd);
}

void orc_tucker_reconstruct(int order, const long *dims, const long *ranks,
                            const double *core, const double *const *U, double *Xhat) {
    long N = prod(dims, order);
#pragma omp parallel for schedule(static)
    for (long e = 0; e < N; ++e) {
        long idx[4] = {0, 0, 0, 0}, rem = e;
        for (int m = 0; m < order; ++m) { idx[m] = rem % dims[m]; rem /= dims[m]; }
        long nc = prod(ranks, order);
        double s = 0.0;
        for (long c = 0; c < nc; ++c) {
            long ci[4] = {0, 0, 0, 0}, rr = c;
            for (int m = 0; m < order; ++m) { ci[m] = rr % ranks[m]; rr /= ranks[m]; }
            double p = core[c];
            for (int m = 0; m < order; ++m) p *= U[m][idx[m] + ci[m] * dims[m]];
            s += p;
        }
        Xhat[e] = s;
    }
}

/* ------------------------------------------------------------------------- */
/* GELU (Hendrycks & Gimpel 2016), exact erf form */
static inline double gelu(double x) { return 0.5 * x * (1.0 + erf(x * 0.70710678118654752440)); }
static inline double gelu_d(double x) {
    return 0.5 * (1.0 + erf(x * 0.70710678118654752440)) + x * 0.39894228040143267794 * exp(-0.5 * x * x);
}

static void dense(const double *in, long batch, long nin, long nout, const double *W, const double *b, double *out) {
    for (long s = 0; s < batch; ++s)
        for (long o = 0; o < nout; ++o) {
            double acc = b[o];
            for (long i = 0; i < nin; ++i) acc += W[o * nin + i] * in[s * nin + i];
            out[s * nout + o] = acc;
        }
}

void orc_mlp_forward(const double *x, long batch, long nin, long width, long nout,
                     const double *W1, const double *b1, const double *W2, const double *b2,
                     const double *W3, const double *b3, double *y) {
    double *h1 = (double *)malloc(sizeof(double) * batch * width);
    double *h2 = (double *)malloc(sizeof(double) * batch * width);
    dense(x, batch, nin, width, W1, b1, h1);
    for (long e = 0; e < batch * width; ++e) h1[e] = gelu(h1[e]);
    dense(h1, batch, width, width, W2, b2, h2);
    for (long e = 0; e < batch * width; ++e) h2[e] = gelu(h2[e]);
    dense(h2, batch, width, nout, W3, b3, y);
    free(h1); free(h2);
}

double orc_mlp_backward(const double *x, const double *t, long batch, long nin, long width, long nout,
                        const double *W1, const double *b1, const double *W2, const double *b2,
                        const double *W3, const double *b3,
                        double *gW1, double *gb1, double *gW2, double *gb2, double *gW3, double *gb3) {
    double *z1 = (double *)malloc(sizeof(double) * batch * width);
    double *h1 = (double *)malloc(sizeof(double) * batch * width);
    double *z2 = (double *)malloc(sizeof(double) * batch * width);
    double *h2 = (double *)malloc(sizeof(double) * batch * width);
    double *y = (double *)malloc(sizeof(double) * batch * nout);
    double *dy = (double *)malloc(sizeof(double) * batch * nout);
    double *d2 = (double *)malloc(sizeof(double) * batch * width);
    double *d1 = (double *)malloc(sizeof(double) * batch * width);
    dense(x, batch, nin, width, W1, b1, z1);
    for (long e = 0; e < batch * width; ++e) h1[e] = gelu(z1[e]);
    dense(h1, batch, width, width, W2, b2, z2);
    for (long e = 0; e < batch * width; ++e) h2[e] = gelu(z2[e]);
    dense(h2, batch, width, nout, W3, b3, y);
    double loss = 0.0, inv = 1.0 / (double)(batch * nout);
    for (long e = 0; e < batch * nout; ++e) {
        double d = y[e] - t[e];
        loss += 0.5 * d * d * inv;
        dy[e] = d * inv;
    }
    memset(gW3, 0, sizeof(double) * nout * width); memset(gb3, 0, sizeof(double) * nout);
    memset(gW2, 0, sizeof(double) * width * width); memset(gb2, 0, sizeof(double) * width);
    memset(gW1, 0, sizeof(double) * width * nin); memset(gb1, 0, sizeof(double) * width);
    for (long s = 0; s < batch; ++s) {
        for (long o = 0; o < nout; ++o) {
            gb3[o] += dy[s * nout + o];
            for (long i = 0; i < width; ++i) gW3[o * width + i] += dy[s * nout + o] * h2[s * width + i];
        }
        for (long i = 0; i < width; ++i) {
            double acc = 0.0;
            for (long o = 0; o < nout; ++o) acc += W3[o * width + i] * dy[s * nout + o];
            d2[s * width + i] = acc * gelu_d(z2[s * width + i]);
        }
        for (long o = 0; o < width; ++o) {
            gb2[o] += d2[s * width + o];
            for (long i = 0; i < width; ++i) gW2[o * width + i] += d2[s * width + o] * h1[s * width + i];
        }
        for (long i = 0; i < width; ++i) {
            double acc = 0.0;
            for (long o = 0; o < width; ++o) acc += W2[o * width + i] * d2[s * width + o];
            d1[s * width + i] = acc * gelu_d(z1[s * width + i]);
        }
        for (long o = 0; o < width; ++o) {
            gb1[o] += d1[s * width + o];
            for (long i = 0; i < nin; ++i) gW1[o * nin + i] += d1[s * width + o] * x[s * nin + i];
        }
    }
    free(z1); free(h1); free(z2); free(h2); free(y); free(dy); free(d2); free(d1);
    return loss;
}

void orc_adam(double *p, const double *g, double *m, double *v, long n, long t,
              double lr, double b1, double b2, double eps) {
    double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
    for (long i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        double mh = m[i] / c1, vh = v[i] / c2;
        p[i] -= lr * mh / (sqrt(vh) + eps);
    }
}
