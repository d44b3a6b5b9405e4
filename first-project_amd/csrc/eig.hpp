// eig.hpp -- eigenvectors of the r largest eigenvalues of a symmetric n x n FP64 matrix
// (n <= 128), in one 256-thread workgroup.  Replaces the cusolverDnDsyevdx call of
// src/tucker_hosvd.cu:43-61 (and src/tucker4_hosvd.cu) for the Tucker Gram eigensolve.
//
// Method (Golub & Van Loan, 4th ed.): Householder tridiagonalisation T = Q^T A Q (8.3.1,
// LAPACK dsytd2 conventions), eigenvalues of T by Sturm-count multisection (8.4.1), their
// eigenvectors by inverse iteration on T with a partially pivoted tridiagonal LU (8.2.2;
// vectors of close eigenvalues re-orthogonalised as LAPACK dstein does), and back
// transformation z = Q y.  Output columns are in ascending eigenvalue order with the sign
// chosen so that the entry of largest magnitude (first one on ties) is positive.
#pragma once
#include "common.hpp"
#include <float.h>
#include <math.h>

namespace tdx {

static constexpr int kEigMaxN = 128;
static constexpr int kEigThreads = 256;
static constexpr int kEigSmall = 10 * kEigMaxN + 64;  // doubles of small vectors
// the large LDS area: the matrix (lda = n + 1), later the packed Householder vectors plus a
// 32-column batch of eigenvectors
// (at least 512 doubles: the bisection counts, 4 ints per thread, also live there)
__host__ __device__ constexpr long eig_area_doubles(int n) {
    const long a = (long)n * (n + 1), b = (long)n * (n - 1) / 2 + 32L * n;
    const long m = a > b ? a : b;
    return m > 512 ? m : 512;
}
// dynamic LDS doubles needed for order n
__host__ __device__ constexpr long eig_lds_doubles(int n) { return kEigSmall + eig_area_doubles(n); }

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// 1/b to full double precision: the hardware estimate refined by two Newton steps
// (r <- r + r (1 - b r)); b is finite and nonzero wherever it is used here
__device__ __forceinline__ double rcp_f64(double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    r = fma(fma(-b, r, 1.0), r, r);
    return r;
}

// number of eigenvalues of the (scaled) tridiagonal (d, e2 = e^2) that are < x
__device__ inline int sturm_count(const double *d, const double *e2, int n, double x, double pivmin) {
    double q = d[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    int c = q < 0.0;
    for (int i = 1; i < n; ++i) {
        q = (d[i] - x) - e2[i - 1] * rcp_f64(q);
        if (fabs(q) <= pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

// the same count at NP points at once (independent recurrences, so their latencies overlap)
template <int NP>
__device__ inline void sturm_count_n(const double *d, const double *e2, int n, const double *x, double pivmin,
                                     int *c) {
    double q[NP];
#pragma unroll
    for (int t = 0; t < NP; ++t) {
        q[t] = d[0] - x[t];
        if (fabs(q[t]) <= pivmin) q[t] = -pivmin;
        c[t] = q[t] < 0.0;
    }
    for (int i = 1; i < n; ++i) {
        const double di = d[i], ei = e2[i - 1];
#pragma unroll
        for (int t = 0; t < NP; ++t) {
            q[t] = (di - x[t]) - ei * rcp_f64(q[t]);
            if (fabs(q[t]) <= pivmin) q[t] = -pivmin;
            c[t] += q[t] < 0.0;
        }
    }
}

// One inverse-iteration solve (T - lam I) x = y for the tridiagonal (d, e), in place on the
// strided vector y (element i at y[i * ys]); u0/u1/u2 are strided scratch rows of the
// pivoted LU.  Returns nothing; the caller normalises.
__device__ inline void tri_solve_shifted(const double *d, const double *e, int n, double lam, double tiny, double *y,
                                         double *u0, double *u1, double *u2, int ys) {
    double alpha = d[0] - lam;
    double beta = (n > 1) ? e[0] : 0.0;
    double ycur = y[0];
    for (int i = 0; i + 1 < n; ++i) {
        const double b = e[i];
        const double a1 = d[i + 1] - lam;
        const double c1 = (i + 2 < n) ? e[i + 1] : 0.0;
        const double ynext = y[(i + 1) * ys];
        double r0, r1, r2, rhs, na, nb, ny;
        if (fabs(alpha) >= fabs(b)) {
            if (alpha == 0.0) alpha = tiny;
            const double l = b * rcp_f64(alpha);
            r0 = alpha; r1 = beta; r2 = 0.0;
            rhs = ycur;
            ny = ynext - l * ycur;
            na = a1 - l * beta;
            nb = c1;
        } else {
            const double l = alpha * rcp_f64(b);
            r0 = b; r1 = a1; r2 = c1;
            rhs = ynext;
            ny = ycur - l * ynext;
            na = beta - l * a1;
            nb = -l * c1;
        }
        u0[i * ys] = rcp_f64(r0);
        u1[i * ys] = r1;
        u2[i * ys] = r2;
        y[i * ys] = rhs;
        ycur = ny;
        alpha = na;
        beta = nb;
    }
    if (fabs(alpha) < tiny) alpha = (alpha < 0.0) ? -tiny : tiny;
    u0[(n - 1) * ys] = rcp_f64(alpha);
    y[(n - 1) * ys] = ycur;
    // back substitution with the upper factor (bandwidth 3)
    double x2 = 0.0, x1 = y[(n - 1) * ys] * u0[(n - 1) * ys];
    y[(n - 1) * ys] = x1;
    for (int i = n - 2; i >= 0; --i) {
        const double x0 = (y[i * ys] - u1[i * ys] * x1 - u2[i * ys] * x2) * u0[i * ys];
        y[i * ys] = x0;
        x2 = x1;
        x1 = x0;
    }
}

// A: n x n symmetric (column-major, lda n) in global memory.  U: n x r FP32 output
// (column-major, ld n).  w (optional): the r eigenvalues, ascending.  Vh: n*n doubles and
// Yg: n*r doubles of global scratch.  Launch with kEigThreads threads and
// eig_lds_doubles(n) * 8 bytes of dynamic LDS.
__device__ inline void sym_top_eigvecs_block(const double *__restrict__ Ag, int n, int r, float *__restrict__ U,
                                             double *__restrict__ w, double *__restrict__ Vh,
                                             double *__restrict__ Yg, double *smem) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lda = n + 1;
    double *dd = smem;                  // diagonal of T
    double *ee = dd + kEigMaxN;         // off-diagonal of T
    double *e2 = ee + kEigMaxN;         // its squares
    double *tt = e2 + kEigMaxN;         // Householder tau
    double *pp = tt + kEigMaxN;         // 2 x kEigMaxN partial products
    double *lam = pp + 2 * kEigMaxN;    // wanted eigenvalues (scaled)
    double *vv = lam + kEigMaxN;        // Householder vector of the current step
    double *ww = vv + kEigMaxN;         // w = p - K v of the current step
    double *red = ww + kEigMaxN;        // 64 reduction slots
    double *A = smem + kEigSmall;
    const long area = eig_area_doubles(n);

    for (int e = tid; e < n * n; e += kEigThreads) A[(e % n) + (e / n) * lda] = Ag[e];
    __syncthreads();

    // ---- 1. tridiagonalisation: for k, H_k = I - tau v v^T zeroes A[k+2:, k] ----
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;
        const double *x = A + (k + 1) + (long)k * lda;
        double s = 0.0;
        for (int i = 1 + tid; i < m; i += kEigThreads) s += x[i] * x[i];
        s = wave_sum_f64(s);
        if (lane == 0) red[wave] = s;
        __syncthreads();
        const double sigma = red[0] + red[1] + red[2] + red[3];
        const double alpha = x[0];
        double tau = 0.0, beta = alpha, sc = 0.0;
        if (sigma > 0.0) {
            const double mu = sqrt(alpha * alpha + sigma);
            beta = (alpha <= 0.0) ? mu : -mu;
            tau = (beta - alpha) / beta;
            sc = 1.0 / (alpha - beta);
        }
        if (tid == 0) {
            dd[k] = A[k + (long)k * lda];
            ee[k] = beta;
            tt[k] = tau;
        }
        if (tid < m) {
            const double v = (tid == 0) ? 1.0 : x[tid] * sc;
            vv[tid] = v;
            Vh[(long)k * n + tid] = v;
        }
        __syncthreads();
        if (tau == 0.0) continue;  // uniform: H_k = I
        // p = tau A22 v, two threads per row, 8 independent partial sums
        const int i = tid & (kEigMaxN - 1), half = tid >> 7;
        double p = 0.0;
        if (i < m) {
            const int mid = (m + 1) >> 1;
            const int j0 = half ? mid : 0, j1 = half ? m : mid;
            const double *arow = A + (k + 1 + i) + (long)(k + 1) * lda;
            double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            int j = j0;
            for (; j + 8 <= j1; j += 8) {
#pragma unroll
                for (int u = 0; u < 8; ++u) acc[u] = fma(arow[(long)(j + u) * lda], vv[j + u], acc[u]);
            }
            for (; j < j1; ++j) acc[0] = fma(arow[(long)j * lda], vv[j], acc[0]);
            p = tau * (((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7])));
        }
        pp[half * kEigMaxN + i] = p;
        const double pv = wave_sum_f64((i < m) ? p * vv[i] : 0.0);
        if (lane == 0) red[4 + wave] = pv;
        __syncthreads();
        const double K = 0.5 * tau * (red[4] + red[5] + red[6] + red[7]);
        if (tid < m) ww[tid] = pp[tid] + pp[kEigMaxN + tid] - K * vv[tid];
        __syncthreads();
        // A22 -= v w^T + w v^T: rows lane, lane + 64; columns j = wave mod 4, 4 at a time
        const int i0 = lane, i1 = lane + 64;
        const bool r0ok = i0 < m, r1ok = i1 < m;
        const double v0 = r0ok ? vv[i0] : 0.0, w0 = r0ok ? ww[i0] : 0.0;
        const double v1 = r1ok ? vv[i1] : 0.0, w1 = r1ok ? ww[i1] : 0.0;
        double *a22 = A + (k + 1) + (long)(k + 1) * lda;
        int j = wave;
        for (; j + 12 < m; j += 16) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int jj = j + 4 * u;
                const double vj = vv[jj], wj = ww[jj];
                double *acol = a22 + (long)jj * lda;
                if (r0ok) acol[i0] -= v0 * wj + w0 * vj;
                if (r1ok) acol[i1] -= v1 * wj + w1 * vj;
            }
        }
        for (; j < m; j += 4) {
            const double vj = vv[j], wj = ww[j];
            double *acol = a22 + (long)j * lda;
            if (r0ok) acol[i0] -= v0 * wj + w0 * vj;
            if (r1ok) acol[i1] -= v1 * wj + w1 * vj;
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (n >= 2) {
            dd[n - 2] = A[(n - 2) + (long)(n - 2) * lda];
            dd[n - 1] = A[(n - 1) + (long)(n - 1) * lda];
            ee[n - 2] = A[(n - 1) + (long)(n - 2) * lda];
        } else {
            dd[0] = A[0];
        }
    }
    __syncthreads();

    // ---- 2. scale T to norm ~1 by a power of two (exact), Gershgorin interval ----
    double tmax = 0.0;
    if (tid < n) {
        tmax = fabs(dd[tid]);
        if (tid + 1 < n) tmax = fmax(tmax, fabs(ee[tid]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tmax = fmax(tmax, __shfl_xor(tmax, off, 64));
    if (lane == 0) red[8 + wave] = tmax;
    __syncthreads();
    tmax = fmax(fmax(red[8], red[9]), fmax(red[10], red[11]));
    int ex = 0;
    if (tmax > 0.0) frexp(tmax, &ex);
    const double scale = ldexp(1.0, -ex);  // T * scale has entries of magnitude < 1
    double glo = 0.0, ghi = 0.0;
    if (tid < n) {
        const double di = dd[tid] * scale;
        const double el = (tid > 0) ? fabs(ee[tid - 1] * scale) : 0.0;
        const double er = (tid + 1 < n) ? fabs(ee[tid] * scale) : 0.0;
        glo = di - el - er;
        ghi = di + el + er;
    }
    __syncthreads();
    if (tid < n) {
        dd[tid] *= scale;
        if (tid + 1 < n) {
            ee[tid] *= scale;
            e2[tid] = ee[tid] * ee[tid];
        }
    }
    const double big = 1e300;
    double mn = (tid < n) ? glo : big, mx = (tid < n) ? ghi : -big;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, off, 64));
        mx = fmax(mx, __shfl_xor(mx, off, 64));
    }
    if (lane == 0) {
        red[12 + wave] = mn;
        red[16 + wave] = mx;
    }
    __syncthreads();
    glo = fmin(fmin(red[12], red[13]), fmin(red[14], red[15]));
    ghi = fmax(fmax(red[16], red[17]), fmax(red[18], red[19]));
    const double tnorm = fmax(fabs(glo), fabs(ghi));
    const double pivmin = DBL_MIN;
    glo -= 4.0 * DBL_EPSILON * tnorm + 2.0 * pivmin;
    ghi += 4.0 * DBL_EPSILON * tnorm + 2.0 * pivmin;

    // ---- 3. multisection for eigenvalues n-r .. n-1: g threads x NP points per eigenvalue,
    //         so each pass narrows the interval (g NP + 1)-fold ----
    constexpr int NP = 4;
    int lg = 0;
    while ((1 << (lg + 1)) < r) ++lg;
    if (r > 1) ++lg;  // 2^lg >= r
    int g = kEigThreads >> lg;
    if (g < 1) g = 1;
    const int npt = g * NP;
    const double inv_div = 1.0 / (double)(npt + 1);
    const int grp = tid / g, t = tid % g;
    const bool active = grp < r;
    const int jwant = n - r + grp;
    int *cnt = reinterpret_cast<int *>(A);  // counts, the matrix area is free now
    double lo = glo, hi = ghi;
    const double atol = DBL_EPSILON * tnorm + 4.0 * pivmin;
    for (int it = 0; it < 200; ++it) {
        const double h = (hi - lo) * inv_div;
        double xs[NP];
        int cs[NP];
#pragma unroll
        for (int u = 0; u < NP; ++u) xs[u] = lo + h * (double)(t * NP + u + 1);
        if (active) {
            sturm_count_n<NP>(dd, e2, n, xs, pivmin, cs);
        } else {
#pragma unroll
            for (int u = 0; u < NP; ++u) cs[u] = 0;
        }
#pragma unroll
        for (int u = 0; u < NP; ++u) cnt[tid * NP + u] = cs[u];
        __syncthreads();
        bool stalled = false;
        if (active) {
            double nlo = lo, nhi = hi;
            for (int u = 0; u < npt; ++u) {
                const double xu = lo + h * (double)(u + 1);
                if (cnt[grp * npt + u] > jwant) {
                    nhi = xu;
                    break;
                }
                nlo = xu;
            }
            stalled = (nlo == lo && nhi == hi);
            lo = nlo;
            hi = nhi;
        }
        const bool done = !active || stalled || (hi - lo) <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + atol;
        if (!__syncthreads_or(!done)) break;
    }
    if (active && t == 0) lam[grp] = 0.5 * (lo + hi);
    __syncthreads();
    if (w)
        for (int c = tid; c < r; c += kEigThreads) w[c] = lam[c] / scale;

    // ---- 4. inverse iteration, batches of up to NB vectors, one thread each ----
    // LDS (the matrix area is free now): y, u0, u1, u2 as [i][c] with c fastest
    const int NB = (area / (4L * n) < 64) ? (int)(area / (4L * n)) : 64;
    const double tiny = DBL_EPSILON * fmax(tnorm, 1.0);
    for (int c0 = 0; c0 < r; c0 += NB) {
        const int nb = (r - c0 < NB) ? r - c0 : NB;
        double *Y = A, *U0 = A + (long)n * nb, *U1 = U0 + (long)n * nb, *U2 = U1 + (long)n * nb;
        if (tid < nb) {
            const int c = tid;
            const double lc = lam[c0 + c];
            for (int i = 0; i < n; ++i) Y[i * nb + c] = 1.0 + 0.25 * (double)((i * 37 + (c0 + c) * 101 + 11) % 61) / 61.0;
            for (int itr = 0; itr < 2; ++itr) {
                tri_solve_shifted(dd, ee, n, lc, tiny, Y + c, U0 + c, U1 + c, U2 + c, nb);
                double mxv = 0.0;
                for (int i = 0; i < n; ++i) mxv = fmax(mxv, fabs(Y[i * nb + c]));
                const double inv = (mxv > 0.0) ? 1.0 / mxv : 1.0;
                for (int i = 0; i < n; ++i) Y[i * nb + c] *= inv;
            }
            double s2 = 0.0;
            for (int i = 0; i < n; ++i) s2 += Y[i * nb + c] * Y[i * nb + c];
            const double inv = 1.0 / sqrt(s2);
            for (int i = 0; i < n; ++i) Yg[(long)(c0 + c) * n + i] = Y[i * nb + c] * inv;
        }
        __syncthreads();
    }
    __threadfence_block();

    // ---- 5. re-orthogonalise vectors of close eigenvalues (clusters), in order ----
    // Y (n x r, column-major) in LDS; cluster = chain of gaps <= 1e-3 * ||T||
    double *Y = A;
    for (int e = tid; e < n * r; e += kEigThreads) Y[e] = Yg[e];
    __syncthreads();
    {
        int cstart = 0;
        for (int c = 1; c < r; ++c) {
            if (lam[c] - lam[c - 1] > 1e-3 * tnorm) {
                cstart = c;
                continue;
            }
            for (int j = cstart; j < c; ++j) {
                double s = 0.0;
                for (int i = tid; i < n; i += kEigThreads) s += Y[(long)j * n + i] * Y[(long)c * n + i];
                s = wave_sum_f64(s);
                if (lane == 0) red[20 + wave] = s;
                __syncthreads();
                s = red[20] + red[21] + red[22] + red[23];
                for (int i = tid; i < n; i += kEigThreads) Y[(long)c * n + i] -= s * Y[(long)j * n + i];
                __syncthreads();
            }
            double s = 0.0;
            for (int i = tid; i < n; i += kEigThreads) s += Y[(long)c * n + i] * Y[(long)c * n + i];
            s = wave_sum_f64(s);
            if (lane == 0) red[24 + wave] = s;
            __syncthreads();
            s = red[24] + red[25] + red[26] + red[27];
            const double inv = 1.0 / sqrt(s);
            for (int i = tid; i < n; i += kEigThreads) Y[(long)c * n + i] *= inv;
            __syncthreads();
        }
    }
    for (int e = tid; e < n * r; e += kEigThreads) Yg[e] = Y[e];
    __syncthreads();

    // ---- 6. back transformation z = H_0 ... H_{n-3} y, column batches, then sign + store ----
    // LDS: packed Householder vectors (column k at koff(k), n-k-1 entries) then the batch
    double *Vp = A;
    const long nv = (n >= 3) ? (long)(n - 2) * (n - 1) - (long)(n - 2) * (n - 3) / 2 : 0;
    for (int k = 0; k + 2 < n; ++k) {
        const long off = (long)k * (n - 1) - (long)k * (k - 1) / 2;
        for (int i = tid; i < n - k - 1; i += kEigThreads) Vp[off + i] = Vh[(long)k * n + i];
    }
    double *Yb = A + nv;
    constexpr int CB = 32;
    constexpr int gb = kEigThreads / CB;  // threads per column
    for (int c0 = 0; c0 < r; c0 += CB) {
        const int nb = (r - c0 < CB) ? r - c0 : CB;
        for (int e = tid; e < n * nb; e += kEigThreads) Yb[e] = Yg[(long)c0 * n + e];
        __syncthreads();
        const int c = tid / gb, tq = tid % gb;
        const bool on = c < nb;  // groups of gb lanes are uniformly on or off
        for (int k = n - 3; k >= 0; --k) {
            const int m = n - k - 1;
            const long off = (long)k * (n - 1) - (long)k * (k - 1) / 2;
            const double tau = tt[k];
            // this thread's rows i = tq + gb u of the reflector (at most kEigMaxN / gb of them)
            constexpr int NR = kEigMaxN / gb;
            double vr[NR], yr[NR];
            double *ycol = Yb + (long)c * n + k + 1;
            double s = 0.0;
#pragma unroll
            for (int u = 0; u < NR; ++u) {
                const int i = tq + gb * u;
                const bool in = on && i < m;
                vr[u] = in ? Vp[off + i] : 0.0;
                yr[u] = in ? ycol[i] : 0.0;
                s = fma(vr[u], yr[u], s);
            }
#pragma unroll
            for (int o = 1; o < gb; o <<= 1) s += __shfl_xor(s, o, 64);
            const double f = tau * s;
#pragma unroll
            for (int u = 0; u < NR; ++u) {
                const int i = tq + gb * u;
                if (on && i < m) ycol[i] = yr[u] - f * vr[u];
            }
            __syncthreads();
        }
        // sign: entry of largest magnitude (lowest index on ties) positive
        if (on) {
            double best = -1.0;
            int bi = n;
            for (int i = tq; i < n; i += gb) {
                const double a = fabs(Yb[(long)c * n + i]);
                if (a > best) {
                    best = a;
                    bi = i;
                }
            }
#pragma unroll
            for (int o = 1; o < gb; o <<= 1) {
                const double ob = __shfl_xor(best, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            const double sg = (Yb[(long)c * n + bi] < 0.0) ? -1.0 : 1.0;
            for (int i = tq; i < n; i += gb) U[(long)(c0 + c) * n + i] = (float)(sg * Yb[(long)c * n + i]);
        }
        __syncthreads();
    }
}

}  // namespace tdx
