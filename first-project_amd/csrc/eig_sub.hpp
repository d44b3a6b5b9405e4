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
// nearly dependent columns.
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
static constexpr int kSubLd = kSubMaxR;       // row stride of the n x r blocks V, Z
static constexpr int kSubHld = kSubMaxR + 1;  // row stride of the r x r blocks H, W, S
static constexpr int kSubMaxIt = 12;

// dynamic LDS doubles of the subspace path for order n: V, Z (n x kSubLd), H, W, S
// (kSubMaxR x kSubHld), cinv, dj (kSubMaxR), part (8 x kSubMaxR), the Jacobi parameters
// (8 x 16) and the reduction slots (64)
__host__ __device__ constexpr long eig_sub_doubles(int n) {
    return 2L * n * kSubLd + 3L * kSubMaxR * kSubHld + 2L * kSubMaxR + 8L * kSubMaxR + 8L * 16 + 64;
}

// sum over the workgroup; every thread gets the total (red: 4 doubles, reused after a barrier)
__device__ inline double sub_blk_sum(double v, double *red) {
    v = wave_sum_f64(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// X (n x r, stride kSubLd, columns >= r zero) <- orthonormal basis of its column span by
// Cholesky QR of the column-normalised block.  Returns the smallest Cholesky pivot (1 for
// orthonormal input, 0 for a zero or numerically dependent column).
__device__ inline double sub_orth_pass(double *X, int n, int r, double *S, double *cinv, double *dj,
                                       double *part) {
    const int tid = threadIdx.x;
    {
        const int c = tid & 31, p = tid >> 5;
        double s = 0.0;
        if (c < r)
            for (int i = p; i < n; i += 8) s = fma(X[i * kSubLd + c], X[i * kSubLd + c], s);
        part[p * 32 + c] = s;
    }
    __syncthreads();
    if (tid < r) {
        double s = 0.0;
        for (int p = 0; p < 8; ++p) s += part[p * 32 + tid];
        cinv[tid] = (s > 0.0) ? 1.0 / sqrt(s) : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < r; ++c)
        if (cinv[c] == 0.0) return 0.0;  // uniform
    // S = Xn^T Xn (lower triangle used), 2 x 2 blocks per thread
    {
        const int a = 2 * (tid >> 4), b = 2 * (tid & 15);
        if (a < r && b <= a) {
            double s00 = 0.0, s01 = 0.0, s10 = 0.0, s11 = 0.0;
            for (int i = 0; i < n; ++i) {
                const double *xr = X + i * kSubLd;
                const double xa0 = xr[a], xa1 = xr[a + 1], xb0 = xr[b], xb1 = xr[b + 1];
                s00 = fma(xa0, xb0, s00);
                s01 = fma(xa0, xb1, s01);
                s10 = fma(xa1, xb0, s10);
                s11 = fma(xa1, xb1, s11);
            }
            const double na0 = cinv[a], nb0 = cinv[b];
            const double na1 = (a + 1 < r) ? cinv[a + 1] : 0.0, nb1 = (b + 1 < r) ? cinv[b + 1] : 0.0;
            S[a * kSubHld + b] = s00 * na0 * nb0;
            if (b + 1 < r) S[a * kSubHld + b + 1] = s01 * na0 * nb1;
            if (a + 1 < r) {
                S[(a + 1) * kSubHld + b] = s10 * na1 * nb0;
                if (b + 1 < r) S[(a + 1) * kSubHld + b + 1] = s11 * na1 * nb1;
            }
        }
    }
    __syncthreads();
    // right-looking Cholesky on the unscaled columns: step j subtracts S_ij S_lj / d_j
    double pmin = 1.0;
    for (int j = 0; j < r; ++j) {
        const double d = S[j * kSubHld + j];
        pmin = fmin(pmin, d);
        if (!(d > 1e-10)) return 0.0;  // uniform: every thread read the same value
        if (tid == 0) dj[j] = d;
        const int m = r - j - 1;
        const double dinv = 1.0 / d;
        for (int e = tid; e < m * m; e += kEigThreads) {
            const int i = j + 1 + e / m, l = j + 1 + e % m;
            if (l <= i) S[i * kSubHld + l] -= S[i * kSubHld + j] * S[l * kSubHld + j] * dinv;
        }
        __syncthreads();
    }
    // L_ij = S_ij / sqrt(d_j), L_jj = sqrt(d_j); keep 1 / L_jj in dj
    for (int e = tid; e < r * r; e += kEigThreads) {
        const int i = e / r, j = e % r;
        if (j < i) S[i * kSubHld + j] *= 1.0 / sqrt(dj[j]);
    }
    __syncthreads();
    if (tid < r) dj[tid] = 1.0 / sqrt(dj[tid]);
    __syncthreads();
    // rows of Q = Xn L^{-T}: forward substitution L q = x per row
    if (tid < n) {
        double q[kSubMaxR];
        double *xr = X + tid * kSubLd;
#pragma unroll
        for (int j = 0; j < kSubMaxR; ++j) {
            double s = 0.0;
            if (j < r) {
                s = xr[j] * cinv[j];
                const double *lr = S + j * kSubHld;
#pragma unroll
                for (int k = 0; k < j; ++k) s = fma(-lr[k], q[k], s);
                s *= dj[j];
            }
            q[j] = s;
        }
#pragma unroll
        for (int j = 0; j < kSubMaxR; ++j)
            if (j < r) xr[j] = q[j];
    }
    __syncthreads();
    return pmin;
}

__device__ inline bool sub_orth(double *X, int n, int r, double *S, double *cinv, double *dj, double *part) {
    double p = sub_orth_pass(X, n, r, S, cinv, dj, part);
    if (p <= 0.0) return false;
    if (p < 1e-2) p = sub_orth_pass(X, n, r, S, cinv, dj, part);  // CholQR2
    return p > 0.0;
}

// Y (n x r) = X (n x r) * M (r x r, stride kSubHld); thread (row, 16-column half)
__device__ inline void sub_rmul(const double *X, const double *M, double *Y, int n, int r) {
    const int i = threadIdx.x & 127, c0 = (threadIdx.x >> 7) * 16;
    if (i < n && c0 < r) {
        double acc[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u] = 0.0;
        const double *xr = X + i * kSubLd;
        for (int k = 0; k < r; ++k) {
            const double x = xr[k];
            const double *mr = M + k * kSubHld + c0;
#pragma unroll
            for (int u = 0; u < 16; ++u) acc[u] = fma(x, mr[u], acc[u]);
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (c0 + u < r) Y[i * kSubLd + c0 + u] = acc[u];
    }
}

// Ag: n x n symmetric (column-major) in global memory.  U0: n x r FP32 start block
// (column-major) or null for the built-in one.  On success writes U (n x r FP32, may alias U0)
// and w (optional) like sym_top_eigvecs_block and returns true; returns false having written
// nothing.  Launch with kEigThreads threads and eig_sub_doubles(n) * 8 bytes of dynamic LDS.
__device__ inline bool sym_top_eigvecs_subspace(const double *Ag, int n, int r, const float *U0, float *U,
                                                double *w, double *smem) {
    const int tid = threadIdx.x;
    double *V = smem;
    double *Z = V + (long)n * kSubLd;
    double *H = Z + (long)n * kSubLd;
    double *Wm = H + kSubMaxR * kSubHld;
    double *S = Wm + kSubMaxR * kSubHld;
    double *cinv = S + kSubMaxR * kSubHld;
    double *dj = cinv + kSubMaxR;
    double *part = dj + kSubMaxR;                // 8 x 32
    double *jpar = part + 8 * kSubMaxR;           // Jacobi: 8 arrays of 16
    double *red = jpar + 8 * 16;                  // reductions
    double *jc = jpar, *js = jpar + 16, *jt = jpar + 32, *japp = jpar + 48, *jaqq = jpar + 64,
           *japq = jpar + 80;
    int *jp = reinterpret_cast<int *>(jpar + 96), *jq = jp + 16;
    int *pos = reinterpret_cast<int *>(jpar + 112);  // 32 ints
    double *sg = red + 8;                            // 32 signs

    // start block; columns r.. stay zero throughout
    for (int e = tid; e < n * kSubLd; e += kEigThreads) {
        const int i = e / kSubLd, c = e % kSubLd;
        double v = 0.0;
        if (c < r) {
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
    double f2 = 0.0;
    for (int e = tid; e < n * n; e += kEigThreads) f2 = fma(Ag[e], Ag[e], f2);
    f2 = sub_blk_sum(f2, red);
    if (!sub_orth(V, n, r, S, cinv, dj, part)) return false;

    double rho = 0.0, hf = 0.0, rho_prev = 0.0;
    bool conv = false;
    for (int it = 0; it < kSubMaxIt; ++it) {
        // Z = A V: thread (row i, 16 columns); A from global (L2-resident), V rows broadcast
        {
            const int i = tid & 127, c0 = (tid >> 7) * 16;
            if (i < n && c0 < r) {
                double acc[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) acc[u] = 0.0;
                const double *acol = Ag + i;
#pragma unroll 4
                for (int k = 0; k < n; ++k) {
                    const double a = acol[(long)k * n];
                    const double *vr = V + k * kSubLd + c0;
#pragma unroll
                    for (int u = 0; u < 16; ++u) acc[u] = fma(a, vr[u], acc[u]);
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) Z[i * kSubLd + c0 + u] = acc[u];
            }
        }
        __syncthreads();
        // S = V^T Z (all r x r entries), 2 x 2 blocks per thread
        {
            const int a = 2 * (tid >> 4), b = 2 * (tid & 15);
            if (a < r && b < r) {
                double s00 = 0.0, s01 = 0.0, s10 = 0.0, s11 = 0.0;
                for (int i = 0; i < n; ++i) {
                    const double va0 = V[i * kSubLd + a], va1 = V[i * kSubLd + a + 1];
                    const double zb0 = Z[i * kSubLd + b], zb1 = Z[i * kSubLd + b + 1];
                    s00 = fma(va0, zb0, s00);
                    s01 = fma(va0, zb1, s01);
                    s10 = fma(va1, zb0, s10);
                    s11 = fma(va1, zb1, s11);
                }
                S[a * kSubHld + b] = s00;
                if (b + 1 < r) S[a * kSubHld + b + 1] = s01;
                if (a + 1 < r) {
                    S[(a + 1) * kSubHld + b] = s10;
                    if (b + 1 < r) S[(a + 1) * kSubHld + b + 1] = s11;
                }
            }
        }
        __syncthreads();
        // rho = ||Z - V S||_F, H = (S + S^T) / 2, ||H||_F
        double rr = 0.0, hh = 0.0;
        {
            const int i = tid & 127, c0 = (tid >> 7) * 16;
            if (i < n && c0 < r) {
                double acc[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) acc[u] = Z[i * kSubLd + c0 + u];
                const double *vr = V + i * kSubLd;
                for (int k = 0; k < r; ++k) {
                    const double v = vr[k];
                    const double *sr = S + k * kSubHld + c0;
#pragma unroll
                    for (int u = 0; u < 16; ++u) acc[u] = fma(-v, sr[u], acc[u]);
                }
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (c0 + u < r) rr = fma(acc[u], acc[u], rr);
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
        if (!sub_orth(Z, n, r, S, cinv, dj, part)) return false;
        double *t = V;
        V = Z;
        Z = t;
    }
    if (!conv || !(hf > 0.0)) return false;

    // ---- Rayleigh-Ritz: H = W diag(theta) W^T by cyclic Jacobi, round-robin pairs ----
    for (int e = tid; e < r * kSubHld; e += kEigThreads) Wm[e] = ((e / kSubHld) == (e % kSubHld)) ? 1.0 : 0.0;
    const int re = r + (r & 1), npair = re >> 1, m = re - 1;
    double off = 0.0;
    bool jconv = false;
    for (int sw = 0; sw < 16; ++sw) {
        for (int k = 0; k < m; ++k) {
            if (tid < npair) {
                int p, q;
                if (tid == 0) {
                    p = re - 1;
                    q = k;
                } else {
                    p = (k + tid) % m;
                    q = (k - tid + m) % m;
                }
                double c = 1.0, s = 0.0, t = 0.0, app = 0.0, aqq = 0.0, apq = 0.0;
                const bool on = p < r && q < r;
                if (on) {
                    app = H[p * kSubHld + p];
                    aqq = H[q * kSubHld + q];
                    apq = H[p * kSubHld + q];
                    if (apq != 0.0) {
                        const double tau = (aqq - app) / (2.0 * apq);
                        t = (fabs(tau) > 1e150) ? 0.5 / tau
                                                : copysign(1.0, tau) / (fabs(tau) + sqrt(fma(tau, tau, 1.0)));
                        c = 1.0 / sqrt(fma(t, t, 1.0));
                        s = t * c;
                    }
                }
                jp[tid] = on ? p : -1;
                jq[tid] = q;
                jc[tid] = c;
                js[tid] = s;
                jt[tid] = t;
                japp[tid] = app;
                jaqq[tid] = aqq;
                japq[tid] = apq;
            }
            __syncthreads();
            // rows p, q of H and columns p, q of W
            for (int e = tid; e < 2 * npair * r; e += kEigThreads) {
                const int which = e / (npair * r), rem = e % (npair * r);
                const int pi = rem / r, j = rem % r;
                const int p = jp[pi];
                if (p < 0) continue;
                const int q = jq[pi];
                const double c = jc[pi], s = js[pi];
                if (which == 0) {
                    const double hp = H[p * kSubHld + j], hq = H[q * kSubHld + j];
                    H[p * kSubHld + j] = c * hp - s * hq;
                    H[q * kSubHld + j] = s * hp + c * hq;
                } else {
                    const double wp = Wm[j * kSubHld + p], wq = Wm[j * kSubHld + q];
                    Wm[j * kSubHld + p] = c * wp - s * wq;
                    Wm[j * kSubHld + q] = s * wp + c * wq;
                }
            }
            __syncthreads();
            // columns p, q of H; the (p, q) block is set exactly
            for (int e = tid; e < npair * r; e += kEigThreads) {
                const int pi = e / r, j = e % r;
                const int p = jp[pi];
                if (p < 0) continue;
                const int q = jq[pi];
                double np_, nq_;
                if (j == p) {
                    np_ = japp[pi] - jt[pi] * japq[pi];
                    nq_ = 0.0;
                } else if (j == q) {
                    np_ = 0.0;
                    nq_ = jaqq[pi] + jt[pi] * japq[pi];
                } else {
                    const double c = jc[pi], s = js[pi];
                    const double hp = H[j * kSubHld + p], hq = H[j * kSubHld + q];
                    np_ = c * hp - s * hq;
                    nq_ = s * hp + c * hq;
                }
                H[j * kSubHld + p] = np_;
                H[j * kSubHld + q] = nq_;
            }
            __syncthreads();
        }
        double o = 0.0;
        for (int e = tid; e < r * r; e += kEigThreads) {
            const int a = e / r, b = e % r;
            if (a != b) o = fma(H[a * kSubHld + b], H[a * kSubHld + b], o);
        }
        off = sqrt(sub_blk_sum(o, red));
        if (off <= 1e-14 * hf) {
            jconv = true;
            break;
        }
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
    sub_rmul(V, Wm, Z, n, r);
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
