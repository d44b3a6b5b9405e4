// eig_sub.hpp -- the r <= 32 largest eigenpairs of a symmetric n x n FP64 matrix (n <= 128)
// by subspace iteration with a Rayleigh-Ritz step, in one 256-thread workgroup, for matrices
// whose top-r eigenvalues stand clear of the rest (the Grams of a Tucker fit: signal
// eigenvalues over a noise floor).  HOOI passes the previous iteration's factor as the start,
// HOSVD a fixed pseudo-random block.
//
// Method (Golub & Van Loan, 4th ed., 8.2.4 orthogonal iteration, 8.1.2 Rayleigh-Ritz, 8.5.2
// parallel cyclic Jacobi):
//   V = orth(V0); repeat { Z = A V; H = V^T Z; rho = ||Z - V H||_F;
//                          stop when rho <= 1e-12 ||H||_F; V = orth(Z) }
//   (H + H^T)/2 = W diag(theta) W^T by cyclic Jacobi in round-robin order; eigenvectors V W.
// orth() is Cholesky QR of the column-normalised block, done twice when the first pass finds
// nearly dependent columns.  Every product (A V, V^T Z, V S, X^T X, X M, V W) runs on FP64
// MFMA (v_mfma_f64_16x16x4f64: exact products, FP64 accumulation), with A held in registers
// for the whole call; the Cholesky factorisation and triangular inverse run in one wave's
// registers; the Jacobi rounds take two barriers each.
//
// The result is accepted only under a certificate that the theta are the r LARGEST
// eigenvalues, valid for any symmetric A: with orthonormal V and residual bound e = rho +
// ||offdiag(H after Jacobi)||_F, r distinct eigenvalues lie within e of the theta (Kahan), so
// every other eigenvalue lambda obeys lambda^2 <= ||A||_F^2 - sum (theta_k - e)^2; the theta are
// the top r when theta_min - e exceeds the square root of that bound (plus 1e-12 ||A||_F^2 for
// rounding).  Otherwise -- slow convergence, rank loss, no certificate -- the caller runs the
// full tridiagonal solver.  Output conventions match it: ascending eigenvalue order, entry of
// largest magnitude (lowest index on ties) positive.
#pragma once
#include "eig.hpp"

namespace tdx {

static constexpr int kSubMaxR = 32;
static constexpr int kSubRows = kEigMaxN;     // V and Z always hold 128 rows (zero past n)
static constexpr int kSubLd = kSubMaxR;       // row stride of the n x r blocks V, Z
static constexpr int kSubHld = kSubMaxR + 1;  // row stride of the r x r blocks H, W, S, M
static constexpr int kSubMaxIt = 12;

// dynamic LDS doubles of the subspace path (any n <= 128): V, Z (128 x kSubLd), H, W, S, M
// (kSubMaxR x kSubHld), 64 reduction slots, signs (32), positions (32 ints) and a flag
__host__ __device__ constexpr long eig_sub_doubles(int) {
    return 2L * kSubRows * kSubLd + 4L * kSubMaxR * kSubHld + 64 + 32 + 16 + 8;
}

typedef double sub_d4 __attribute__((ext_vector_type(4)));
// D += A B on a 16 x 16 tile: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] of the
// 4-deep slice; register v of lane l holds D[l / 16 + 4 v][l % 16]
__device__ __forceinline__ sub_d4 sub_mfma(double a, double b, sub_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// x of lane j (j wave-uniform)
__device__ __forceinline__ double sub_rl(double x, int j) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, j), hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), j);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// 1/sqrt(a) to full double precision: the hardware estimate refined by two Newton steps
// (y <- y (3 - a y^2) / 2); a is finite and positive wherever it is used here
__device__ __forceinline__ double rsq_f64(double a) {
    double y = __builtin_amdgcn_rsq(a);
    y = y * fma(-0.5 * a * y, y, 1.5);
    y = y * fma(-0.5 * a * y, y, 1.5);
    return y;
}

// sum over the workgroup; every thread gets the total (red: 4 doubles, reused after a barrier)
__device__ inline double sub_blk_sum(double v, double *red) {
    v = wave_sum_f64(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Out (32 x 32, stride kSubHld) = X^T Y over the 128 rows of X, Y (stride kSubLd); wave w
// computes tile (w / 2, w % 2)
__device__ inline void sub_gram(const double *X, const double *Y, double *Out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lk = lane >> 4;
    const int tr = wv >> 1, tc = wv & 1;
    sub_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
    for (int ks = 0; ks < kSubRows / 4; ++ks) {
        const int k = 4 * ks + lk;
        acc = sub_mfma(X[k * kSubLd + 16 * tr + lr], Y[k * kSubLd + 16 * tc + lr], acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) Out[(16 * tr + lk + 4 * v) * kSubHld + 16 * tc + lr] = acc[v];
}

// acc = rows 32 w .. 32 w + 31 of X (128 x 32, stride kSubLd) times M (32 x 32, stride kSubHld)
__device__ inline void sub_xm(const double *X, const double *M, sub_d4 (&acc)[2][2]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = sub_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < kSubMaxR / 4; ++ks) {
        const int k = 4 * ks + lk;
        const double a0 = X[(32 * wv + lr) * kSubLd + k], a1 = X[(32 * wv + 16 + lr) * kSubLd + k];
        const double b0 = M[k * kSubHld + lr], b1 = M[k * kSubHld + 16 + lr];
        acc[0][0] = sub_mfma(a0, b0, acc[0][0]);
        acc[0][1] = sub_mfma(a0, b1, acc[0][1]);
        acc[1][0] = sub_mfma(a1, b0, acc[1][0]);
        acc[1][1] = sub_mfma(a1, b1, acc[1][1]);
    }
}

// stores the rows < n of a sub_xm / product result into Y (stride kSubLd)
__device__ inline void sub_store_rows(const sub_d4 (&acc)[2][2], double *Y, int n) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 32 * wv + 16 * rt + lk + 4 * v;
            if (i < n) {
                Y[i * kSubLd + lr] = acc[rt][0][v];
                Y[i * kSubLd + 16 + lr] = acc[rt][1][v];
            }
        }
}

// X (rows < n, columns < r; zero elsewhere) <- orthonormal basis of its column span by
// Cholesky QR of the column-normalised block: S = X^T X, its normalised Cholesky factor L and
// L^{-1} in wave 0's registers (lane l holds row l), X <- X diag(1/sqrt(S_jj)) L^{-T}.
// Returns the smallest Cholesky pivot (1 for orthonormal input; 0 when a column is zero or
// numerically dependent: the call then leaves X unchanged).
__device__ inline double sub_orth_pass(double *X, int n, int r, double *S, double *M, double *flag) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    sub_gram(X, X, S);
    __syncthreads();
    if (wv == 0) {
        const int l = lane;
        double a[kSubMaxR];
#pragma unroll
        for (int j = 0; j < kSubMaxR; ++j) a[j] = (l < r && j < r) ? S[l * kSubHld + j] : 0.0;
        const double dg = (l < r) ? S[l * kSubHld + l] : 1.0;
        const double cl = (dg > 0.0) ? 1.0 / sqrt(dg) : 0.0;
        bool bad = false;
#pragma unroll
        for (int j = 0; j < kSubMaxR; ++j) {
            const double cj = sub_rl(cl, j);
            bad |= (j < r) && !(cj > 0.0);
            a[j] *= cl * cj;
        }
        double pmin = 1.0;
        // right-looking Cholesky of the normalised Gram, whole symmetric matrix kept
#pragma unroll
        for (int j = 0; j < kSubMaxR; ++j) {
            if (j < r && !bad) {
                const double d = sub_rl(a[j], j);
                pmin = fmin(pmin, d);
                if (!(d > 1e-10)) {
                    bad = true;
                } else {
                    const double inv = 1.0 / sqrt(d);
                    const double lj = a[j] * inv;  // L[l][j] (for l == j: sqrt(d))
                    a[j] = lj;
#pragma unroll
                    for (int m = j + 1; m < kSubMaxR; ++m) a[m] = fma(-lj, sub_rl(lj, m), a[m]);
                }
            }
        }
        // lane c < r: column c of L^{-1} by forward substitution, then row c of
        // M = diag(cl) L^{-T}: M[c][i] = cl_c * Linv[i][c]
        if (!bad) {
            double y[kSubMaxR];
#pragma unroll
            for (int i = 0; i < kSubMaxR; ++i) {
                double sacc = (i == l) ? 1.0 : 0.0;
#pragma unroll
                for (int j = 0; j < i; ++j) sacc = fma(-sub_rl(a[j], i), y[j], sacc);
                y[i] = (i < r) ? sacc / sub_rl(a[i], i) : 0.0;
            }
            if (l < kSubMaxR)
#pragma unroll
                for (int i = 0; i < kSubMaxR; ++i) M[l * kSubHld + i] = (l < r) ? cl * y[i] : 0.0;
        }
        if (l == 0) flag[0] = bad ? 0.0 : pmin;
    }
    __syncthreads();
    const double pm = flag[0];
    if (!(pm > 0.0)) return 0.0;  // uniform
    sub_d4 acc[2][2];
    sub_xm(X, M, acc);
    __syncthreads();  // every wave has read X
    sub_store_rows(acc, X, n);
    __syncthreads();
    return pm;
}

__device__ inline bool sub_orth(double *X, int n, int r, double *S, double *M, double *flag) {
    double p = sub_orth_pass(X, n, r, S, M, flag);
    if (p <= 0.0) return false;
    if (p < 1e-2) p = sub_orth_pass(X, n, r, S, M, flag);  // CholQR2
    return p > 0.0;
}

// Ag: n x n symmetric (column-major) in global memory.  U0: n x r FP32 start block
// (column-major) or null for the built-in one.  On success writes U (n x r FP32, may alias U0)
// and w (optional) like sym_top_eigvecs_block and returns true; returns false having written
// nothing.  Launch with kEigThreads threads and eig_sub_doubles(n) * 8 bytes of dynamic LDS.
// rotate = false (HOOI before its last iteration, where only the span of U matters: every
// later product contracts U U^T): when the span is certified cheaply (below) U is the
// orthonormal basis V itself, in no particular order and without the Rayleigh-Ritz rotation;
// otherwise the full path runs.
__device__ inline bool sym_top_eigvecs_subspace(const double *Ag, int n, int r, const float *U0, float *U,
                                                double *w, double *smem, bool rotate = true) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
    double *V = smem;
    double *Z = V + kSubRows * kSubLd;
    double *H = Z + kSubRows * kSubLd;
    double *Wm = H + kSubMaxR * kSubHld;
    double *S = Wm + kSubMaxR * kSubHld;
    double *M = S + kSubMaxR * kSubHld;
    double *red = M + kSubMaxR * kSubHld;  // 64 slots
    double *sg = red + 64;                 // 32 signs
    int *pos = reinterpret_cast<int *>(sg + 32);  // 32 ints
    double *flag = sg + 48;

    // start block; rows n.. and columns r.. stay zero throughout
    for (int e = tid; e < kSubRows * kSubLd; e += kEigThreads) {
        const int i = e / kSubLd, c = e % kSubLd;
        double v = 0.0;
        if (c < r && i < n) {
            if (U0) {
                v = (double)U0[(long)c * n + i];
            } else {
                // integer mix of (i, c) to a uniform value in [-1, 1); the blocks it gives for
                // n <= 128, r <= 32 have column-normalised condition numbers below 600
                unsigned x = ((unsigned)i * 0x9E3779B1u) ^ ((unsigned)(c + 1) * 0x85EBCA77u);
                x ^= x >> 16;
                x *= 0x7FEB352Du;
                x ^= x >> 15;
                x *= 0x846CA68Bu;
                x ^= x >> 16;
                v = (double)(x >> 8) * (2.0 / 16777216.0) - 1.0;
            }
        }
        V[e] = v;
        Z[e] = 0.0;
    }
    // A in registers as MFMA A operands: rows 32 wv + 16 rt + lr, columns 4 ks + lk
    double gA[2][kSubRows / 4];
    double f2 = 0.0;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ks = 0; ks < kSubRows / 4; ++ks) {
            const int i = 32 * wv + 16 * rt + lr, k = 4 * ks + lk;
            const double g = (i < n && k < n) ? Ag[i + (long)k * n] : 0.0;
            gA[rt][ks] = g;
            f2 = fma(g, g, f2);
        }
    f2 = sub_blk_sum(f2, red);  // (its barriers also publish the start block)
    if (!sub_orth(V, n, r, S, M, flag)) return false;

    double rho = 0.0, hf = 0.0, rho_prev = 0.0;
    bool conv = false;
    for (int it = 0; it < kSubMaxIt; ++it) {
        // Z = A V
        {
            sub_d4 acc[2][2];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = sub_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < kSubRows / 4; ++ks) {
                const double *vr = V + (4 * ks + lk) * kSubLd + lr;
                const double b0 = vr[0], b1 = vr[16];
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                    acc[rt][0] = sub_mfma(gA[rt][ks], b0, acc[rt][0]);
                    acc[rt][1] = sub_mfma(gA[rt][ks], b1, acc[rt][1]);
                }
            }
            sub_store_rows(acc, Z, n);
        }
        __syncthreads();
        // S = V^T Z
        sub_gram(V, Z, S);
        __syncthreads();
        // rho = ||Z - V S||_F, H = (S + S^T) / 2, ||H||_F
        double rr = 0.0, hh = 0.0;
        {
            sub_d4 acc[2][2];
            sub_xm(V, S, acc);
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = 32 * wv + 16 * rt + lk + 4 * v, c = 16 * ct + lr;
                        if (i < n && c < r) {
                            const double d = Z[i * kSubLd + c] - acc[rt][ct][v];
                            rr = fma(d, d, rr);
                        }
                    }
        }
        for (int e = tid; e < r * r; e += kEigThreads) {
            const int a = e / r, b = e % r;
            const double h = 0.5 * (S[a * kSubHld + b] + S[b * kSubHld + a]);
            H[a * kSubHld + b] = h;
            hh = fma(h, h, hh);
        }
        rho = sqrt(sub_blk_sum(rr, red));
        hf = sqrt(sub_blk_sum(hh, red + 4));
        if (rho <= 1e-12 * hf) {
            conv = true;
            break;
        }
        if (it >= 2 && rho > 0.25 * rho_prev) return false;  // too slow for this path
        rho_prev = rho;
        if (!sub_orth(Z, n, r, S, M, flag)) return false;
        double *t = V;
        V = Z;
        Z = t;
    }
    if (!conv || !(hf > 0.0)) return false;

    if (!rotate) {
        // Span certificate without the eigenvalues of H: with orthonormal V and residual rho,
        // r eigenvalues lambda_k of A lie within rho of the eigenvalues theta_k of H (Kahan).  If
        // every theta_k > t = rho + b with b^2 = ||A||_F^2 - sum (theta_k - rho)^2 (+ rounding),
        // then lambda_k > b, and every other eigenvalue has lambda^2 <= ||A||_F^2 -
        // sum lambda_k^2 <= b^2: the span of V is the top-r invariant subspace.  sum (theta_k -
        // rho)^2 = ||H||_F^2 - 2 rho tr(H) + r rho^2, and theta_min > t holds exactly when
        // H - t I has a Cholesky factorisation with positive pivots.
        double tr = 0.0;
        for (int c = 0; c < r; ++c) tr += H[c * kSubHld + c];
        const double ss = hf * hf - 2.0 * rho * tr + (double)r * rho * rho;
        const double b2 = fmax(f2 - ss, 0.0) + 1e-12 * f2;
        const double t = rho + sqrt(b2);
        for (int e = tid; e < r * r; e += kEigThreads) {
            const int a = e / r, bb = e % r;
            S[a * kSubHld + bb] = H[a * kSubHld + bb] - ((a == bb) ? t : 0.0);
        }
        if (tid == 0) *flag = 1.0;
        __syncthreads();
        // right-looking Cholesky of the r x r block, one barrier per pivot
        for (int k = 0; k < r; ++k) {
            const double d = S[k * kSubHld + k];
            if (!(d > 0.0)) {
                if (tid == 0) *flag = 0.0;
                break;  // uniform: every thread read the same pivot
            }
            const double dinv = 1.0 / d;
            const int m2 = r - k - 1;
            for (int e = tid; e < m2 * m2; e += kEigThreads) {
                const int i = k + 1 + e / m2, j = k + 1 + e % m2;
                S[i * kSubHld + j] = fma(-S[i * kSubHld + k] * dinv, S[k * kSubHld + j], S[i * kSubHld + j]);
            }
            __syncthreads();
        }
        __syncthreads();
        if (*flag > 0.0) {
            for (int e = tid; e < n * r; e += kEigThreads) {
                const int i = e % n, c = e / n;
                U[(long)c * n + i] = (float)V[i * kSubLd + c];
            }
            return true;
        }
        __syncthreads();
    }

    // ---- Rayleigh-Ritz: H = W diag(theta) W^T by cyclic Jacobi, round-robin pairs ----
    // Thread (group g = tid / 32, index j = tid % 32) handles pairs g and g + 8: it computes
    // their rotations itself (the 32 threads of a group share a wave, so all of them read the
    // pair's entries before any of them writes), rotates rows p, q of H in column j and
    // columns p, q of W in row j, then (after a barrier) columns p, q of H in row j.
    for (int e = tid; e < kSubMaxR * kSubHld; e += kEigThreads) Wm[e] = ((e / kSubHld) == (e % kSubHld)) ? 1.0 : 0.0;
    __syncthreads();
    const int re = r + (r & 1), npair = re >> 1, m = re - 1;
    const int jg = tid >> 5, jj = tid & 31;
    double off = 0.0;
    bool jconv = (m <= 0);
    for (int sw = 0; sw < 16 && !jconv; ++sw) {
        for (int k = 0; k < m; ++k) {
            int pp[2], qq[2];
            double c_[2], s_[2], t_[2], app_[2], aqq_[2], apq_[2];
            bool on[2];
#pragma unroll
            for (int h2 = 0; h2 < 2; ++h2) {
                const int pi = jg + 8 * h2;
                int p, q;
                if (pi == 0) {
                    p = re - 1;
                    q = k;
                } else {
                    p = k + pi;
                    if (p >= m) p -= m;
                    q = k - pi + m;
                    if (q >= m) q -= m;
                }
                on[h2] = pi < npair && p < r && q < r;
                pp[h2] = p;
                qq[h2] = q;
                double c = 1.0, s = 0.0, t = 0.0, app = 0.0, aqq = 0.0, apq = 0.0;
                if (on[h2]) {
                    app = H[p * kSubHld + p];
                    aqq = H[q * kSubHld + q];
                    apq = H[p * kSubHld + q];
                    if (apq != 0.0) {
                        // t = sign(tau) / (|tau| + sqrt(tau^2 + 1)), c = 1 / sqrt(t^2 + 1), with
                        // the hardware reciprocal / reciprocal-root estimates refined by two
                        // Newton steps (full FP64 precision; c^2 + s^2 = 1 to rounding)
                        const double tau = (aqq - app) * rcp_f64(2.0 * apq);
                        if (fabs(tau) > 1e150) {
                            t = 0.5 * rcp_f64(tau);
                        } else {
                            const double a2 = fma(tau, tau, 1.0);
                            const double y = rsq_f64(a2);
                            t = copysign(1.0, tau) * rcp_f64(fma(a2, y, fabs(tau)));
                        }
                        c = rsq_f64(fma(t, t, 1.0));
                        s = t * c;
                    }
                }
                c_[h2] = c;
                s_[h2] = s;
                t_[h2] = t;
                app_[h2] = app;
                aqq_[h2] = aqq;
                apq_[h2] = apq;
            }
            // rows p, q of H (column jj) and columns p, q of W (row jj)
#pragma unroll
            for (int h2 = 0; h2 < 2; ++h2) {
                if (!on[h2] || jj >= r) continue;
                const int p = pp[h2], q = qq[h2];
                const double c = c_[h2], s = s_[h2];
                const double hp = H[p * kSubHld + jj], hq = H[q * kSubHld + jj];
                const double wp = Wm[jj * kSubHld + p], wq = Wm[jj * kSubHld + q];
                H[p * kSubHld + jj] = c * hp - s * hq;
                H[q * kSubHld + jj] = s * hp + c * hq;
                Wm[jj * kSubHld + p] = c * wp - s * wq;
                Wm[jj * kSubHld + q] = s * wp + c * wq;
            }
            __syncthreads();
            // columns p, q of H (row jj); the (p, q) block is set exactly
#pragma unroll
            for (int h2 = 0; h2 < 2; ++h2) {
                if (!on[h2] || jj >= r) continue;
                const int p = pp[h2], q = qq[h2];
                double np_, nq_;
                if (jj == p) {
                    np_ = app_[h2] - t_[h2] * apq_[h2];
                    nq_ = 0.0;
                } else if (jj == q) {
                    np_ = 0.0;
                    nq_ = aqq_[h2] + t_[h2] * apq_[h2];
                } else {
                    const double c = c_[h2], s = s_[h2];
                    const double hp = H[jj * kSubHld + p], hq = H[jj * kSubHld + q];
                    np_ = c * hp - s * hq;
                    nq_ = s * hp + c * hq;
                }
                H[jj * kSubHld + p] = np_;
                H[jj * kSubHld + q] = nq_;
            }
            __syncthreads();
        }
        double o = 0.0;
        for (int e = tid; e < r * r; e += kEigThreads) {
            const int a = e / r, b = e % r;
            if (a != b) o = fma(H[a * kSubHld + b], H[a * kSubHld + b], o);
        }
        off = sqrt(sub_blk_sum(o, red));
        jconv = off <= 1e-14 * hf;
    }
    if (!jconv) return false;

    // ---- certificate that the theta are the r largest eigenvalues ----
    const double eb = rho + off;
    double tmin = 1e300, ssq = 0.0;
    for (int c = 0; c < r; ++c) {
        const double th = H[c * kSubHld + c];
        tmin = fmin(tmin, th);
        const double d = fmax(th - eb, 0.0);
        ssq = fma(d, d, ssq);
    }
    const double rest = fmax(f2 - ssq, 0.0) + 1e-12 * f2;
    if (!(tmin - eb > sqrt(rest))) return false;

    // ---- eigenvectors Y = V W (into Z), ascending order, sign, FP32 ----
    {
        sub_d4 acc[2][2];
        sub_xm(V, Wm, acc);
        sub_store_rows(acc, Z, n);
    }
    if (tid < r) {
        const double th = H[tid * kSubHld + tid];
        int rank = 0;
        for (int k = 0; k < r; ++k) {
            const double tk = H[k * kSubHld + k];
            rank += (tk < th) || (tk == th && k < tid);
        }
        pos[tid] = rank;
        if (w) w[rank] = th;
    }
    __syncthreads();
    {
        const int c = tid >> 3, part8 = tid & 7;  // 8 lanes of one wave per column
        double best = -1.0;
        int bi = n;
        if (c < r)
            for (int i = part8; i < n; i += 8) {
                const double a = fabs(Z[i * kSubLd + c]);
                if (a > best) {
                    best = a;
                    bi = i;
                }
            }
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            const double ob = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (c < r && part8 == 0) sg[c] = (Z[bi * kSubLd + c] < 0.0) ? -1.0 : 1.0;
    }
    __syncthreads();
    for (int e = tid; e < n * r; e += kEigThreads) {
        const int i = e % n, c = e / n;
        U[(long)pos[c] * n + i] = (float)(sg[c] * Z[i * kSubLd + c]);
    }
    return true;
}

}  // namespace tdx
