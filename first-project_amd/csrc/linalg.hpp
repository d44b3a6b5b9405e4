// linalg.hpp -- single-workgroup FP64 dense kernels shared by the CP and Tucker paths.
#pragma once
#include "common.hpp"
#include <math.h>

namespace tdx {

// ---------------------------------------------------------------------------
// block reduction helper (FP64)
__device__ inline double block_sum(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Cyclic Jacobi with round-robin (tournament) pair ordering, one workgroup, FP64.
// A (n x n, lda n) is diagonalised in place; Q accumulates the eigenvectors.
// Golub & Van Loan, section 8.5 (parallel Jacobi ordering, 8.5.4).
__device__ inline void jacobi_eig_block(double *A, double *Q, int n, double *cs, double *sh) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int ne = n + (n & 1);  // even number of players; index n is a dummy if n is odd
    for (int e = tid; e < n * n; e += nt) Q[e] = ((e % n) == (e / n)) ? 1.0 : 0.0;
    __syncthreads();
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int e = tid; e < n * n; e += nt) {
            double v = A[e] * A[e];
            tot += v;
            if ((e % n) != (e / n)) off += v;
        }
        off = block_sum(off, sh);
        tot = block_sum(tot, sh);
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int step = 0; step < ne - 1; ++step) {
            // pair k: positions k and ne-1-k; position 0 holds player 0, position p>0
            // holds player ((p - 1 + step) % (ne - 1)) + 1
            for (int k = tid; k < ne / 2; k += nt) {
                int pa = k, pb = ne - 1 - k;
                int ia = (pa == 0) ? 0 : ((pa - 1 + step) % (ne - 1)) + 1;
                int ib = ((pb - 1 + step) % (ne - 1)) + 1;
                int p = ia < ib ? ia : ib, q = ia < ib ? ib : ia;
                double c = 1.0, s = 0.0;
                if (q < n) {
                    double apq = A[p + q * n];
                    if (apq != 0.0) {
                        double app = A[p + p * n], aqq = A[q + q * n];
                        double tau = (aqq - app) / (2.0 * apq);
                        double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                    }
                }
                cs[4 * k + 0] = c;
                cs[4 * k + 1] = s;
                cs[4 * k + 2] = (double)p;
                cs[4 * k + 3] = (double)q;
            }
            __syncthreads();
            // columns: A := A J, Q := Q J
            for (int e = tid; e < (ne / 2) * n; e += nt) {
                int k = e / n, row = e % n;
                int p = (int)cs[4 * k + 2], q = (int)cs[4 * k + 3];
                if (q >= n) continue;
                double c = cs[4 * k], s = cs[4 * k + 1];
                double akp = A[row + p * n], akq = A[row + q * n];
                A[row + p * n] = c * akp - s * akq;
                A[row + q * n] = s * akp + c * akq;
                double qkp = Q[row + p * n], qkq = Q[row + q * n];
                Q[row + p * n] = c * qkp - s * qkq;
                Q[row + q * n] = s * qkp + c * qkq;
            }
            __syncthreads();
            // rows: A := J^T A
            for (int e = tid; e < (ne / 2) * n; e += nt) {
                int k = e / n, col = e % n;
                int p = (int)cs[4 * k + 2], q = (int)cs[4 * k + 3];
                if (q >= n) continue;
                double c = cs[4 * k], s = cs[4 * k + 1];
                double apk = A[p + col * n], aqk = A[q + col * n];
                A[p + col * n] = c * apk - s * aqk;
                A[q + col * n] = s * apk + c * aqk;
            }
            __syncthreads();
        }
    }
}

// Vinv = pseudo-inverse of the symmetric R x R matrix in L (destroyed) from its Jacobi
// eigendecomposition; eigenvalues below 1e-12 of the largest are dropped.  W receives the
// eigenvectors (R*R doubles), cs needs 4*(R/2+1) doubles, sh blockDim doubles.
__device__ inline void pinv_jacobi(double *L, double *W, double *Vinv, int R, double *cs, double *sh) {
    const int tid = threadIdx.x, nt = blockDim.x, RR = R * R;
    jacobi_eig_block(L, W, R, cs, sh);
    double wmax = 0.0;
    for (int a = 0; a < R; ++a) wmax = fmax(wmax, fabs(L[a + a * R]));
    for (int e = tid; e < RR; e += nt) {
        const int a = e % R, b = e / R;
        double s = 0.0;
        for (int k = 0; k < R; ++k) {
            const double lam = L[k + k * R];
            if (fabs(lam) > 1e-12 * wmax) s += W[a + k * R] * W[b + k * R] / lam;
        }
        Vinv[e] = s;
    }
    __syncthreads();
}

// Vinv = V^{-1} for a symmetric R x R V (all arrays column-major, R*R doubles, any memory
// space visible to the whole workgroup; V is preserved unless L is V, in which case the
// fallback below works in V).  If V is not numerically SPD (a
// relative pivot of the sweep operator below 1e-13), a pseudo-inverse from the Jacobi
// eigendecomposition (eigenvalues below 1e-12 * max dropped) is returned instead and the
// function returns 1 (in every thread).  cs needs 4*(R/2+1) doubles, sh blockDim doubles.
// MAXEPT bounds the elements of V each thread holds: ceil(R^2 / blockDim) <= MAXEPT.
template <int MAXEPT = 16>
__device__ inline int block_spd_inverse(const double *V, double *L, double *W, double *Vinv, int R, double *cs,
                                        double *sh, int *fail) {
    const int tid = threadIdx.x, nt = blockDim.x, RR = R * R;
    // Symmetric sweep operator (Goodnight 1979) on a copy of V, pivots k = 0..R-1 in order:
    //   d = A[k][k];  A[i][j] -= A[i][k] A[k][j] / d  (i, j != k);
    //   A[i][k] = A[k][i] = A[i][k] / d;  A[k][k] = -1 / d.
    // After all R sweeps A = -V^{-1}.  The pivots are the Cholesky pivots L_kk^2; V is treated
    // as numerically SPD when every pivot exceeds 1e-13 of its original diagonal entry (a
    // smaller one means column k is dependent on the earlier ones to rounding).
    // Each thread keeps its elements e = tid + nt u of A in registers; only the pivot column
    // goes through LDS, alternating between two halves of W so that one barrier per pivot
    // suffices (a thread writes pivot k+1's column only after every thread has passed pivot
    // k's barrier, i.e. finished reading pivot k-1's).  W needs 2R doubles.
    // Every one of the MAXEPT register slots is processed (no runtime bound on the loops:
    // guards on a runtime count became a branch per slot); slots past R^2 hold zeros at
    // index (0, 0) and only the pivot-column stores and the final write look at ok.
    constexpr int kMaxEpt = MAXEPT;
    double a[kMaxEpt];
    int ii[kMaxEpt], jj[kMaxEpt];
    bool ok[kMaxEpt];
#pragma unroll
    for (int u = 0; u < kMaxEpt; ++u) {
        const int e = tid + nt * u;
        ok[u] = e < RR;
        ii[u] = ok[u] ? e % R : 0;
        jj[u] = ok[u] ? e / R : 0;
        a[u] = ok[u] ? V[e] : 0.0;
    }
    if (tid == 0) *fail = 0;
    __syncthreads();
    // when R divides the block size a thread's elements share one row index
    const bool one_row = (nt % R) == 0;
    for (int k = 0; k < R; ++k) {
        double *Wk = W + (k & 1) * R;
#pragma unroll
        for (int u = 0; u < kMaxEpt; ++u)
            if (ok[u] && jj[u] == k) Wk[ii[u]] = a[u];
        __syncthreads();
        const double d = Wk[k];
        if (!(d > 1e-13 * V[k + k * R]) || !isfinite(d)) {
            if (tid == 0) *fail = 1;
            break;  // uniform: every thread read the same pivot
        }
        const double dinv = 1.0 / d;
        // all reads of the pivot column first, then a branch-free update
        double wi[kMaxEpt], wj[kMaxEpt];
        if (one_row) {
            const double w0 = Wk[ii[0]];
#pragma unroll
            for (int u = 0; u < kMaxEpt; ++u) {
                wi[u] = w0;
                wj[u] = Wk[jj[u]];
            }
        } else {
#pragma unroll
            for (int u = 0; u < kMaxEpt; ++u) {
                wi[u] = Wk[ii[u]];
                wj[u] = Wk[jj[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < kMaxEpt; ++u) {
            const bool ki = ii[u] == k, kj = jj[u] == k;
            double v = fma(-wi[u] * dinv, wj[u], a[u]);
            v = kj ? wi[u] * dinv : v;
            v = ki ? wj[u] * dinv : v;
            v = (ki && kj) ? -dinv : v;
            a[u] = v;
        }
    }
    __syncthreads();
    if (!*fail) {
#pragma unroll
        for (int u = 0; u < kMaxEpt; ++u)
            if (ok[u]) Vinv[tid + nt * u] = -a[u];
        __syncthreads();
        return 0;
    }
    if (L != V)
        for (int e = tid; e < RR; e += nt) L[e] = V[e];
    __syncthreads();
    pinv_jacobi(L, W, Vinv, R, cs, sh);
    return 1;
}

// Vinv = V^{-1} for a symmetric R x R V (R <= 4 NS <= 64) by the symmetric sweep operator
// of block_spd_inverse (same pivots, same operations per element, so the same result bit
// for bit), for a workgroup of exactly 256 threads.  vat(i, c) returns V[i][c] (read once
// per element, at the start); diag[k] = V[k][k] (any memory visible to the workgroup).
// Layout: lane i of wave w holds row i (rows >= R are zero) of the columns c = w + 4 s in
// register slot s, so the slot holding pivot column k is slot k/4 of wave k%4 (a uniform
// register index, s_set_gpr_idx) and only that wave stores the column (one 8-byte LDS write
// per lane, a uniform branch).  The column
// goes to an LDS half Wb[(k & 1) * 64 ..], permuted so that the values a wave needs (rows
// w, w+4, ...) are contiguous: the entry of row c at (c % 4) * 16 + c / 4 (16-byte broadcast
// reads).  One barrier per pivot (the halves alternate).  Returns 1 (every thread) when a
// relative pivot is below 1e-13 or not finite; Vinv is then not written.  Wb: 128 doubles.
// diag (R doubles, LDS) receives V[k][k] from the lanes that hold it.
template <int NS, typename VAt>
__device__ inline int sweep_inverse_rows(VAt &vat, double *diag, double *Wb, double *Vinv, int R, int *fail) {
    static_assert(NS == 4 || NS == 8 || NS == 16, "NS: slots per lane");
    const int tid = threadIdx.x, i = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    double a[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {  // every load issued, then one wait (see VGather)
        const int c = w + 4 * s;
        vat.load(s, i, c, i < R && c < R);
    }
    vat.finish(a);
#pragma unroll
    for (int s = 0; s < NS; ++s)
        if (w + 4 * s == i && i < R) diag[i] = a[s];
    const int pi = (i & 3) * 16 + (i >> 2);  // this lane's row in the permuted column
    if (tid == 0) *fail = 0;
    __syncthreads();
    int bad = 0;
    for (int k = 0; k < 4 * NS; ++k) {
        if (k >= R) break;
        double *Wk = Wb + (k & 1) * 64;
        if (w == (k & 3)) Wk[pi] = a[k >> 2];
        __syncthreads();
        const double d = Wk[(k & 3) * 16 + (k >> 2)];
        if (!(d > 1e-13 * diag[k]) || !isfinite(d)) {
            bad = 1;
            break;  // uniform: every thread read the same pivot
        }
        const double dinv = 1.0 / d;
        const double t = Wk[pi] * dinv;
        const bool rk = i == k;
        const double alpha = rk ? 0.0 : 1.0, beta = rk ? dinv : -t;
        const double *wr = Wk + w * 16;  // A[w + 4 s][k], s = 0 .. NS-1
#pragma unroll
        for (int s = 0; s < NS; ++s) a[s] = fma(beta, wr[s], alpha * a[s]);
        if (w == (k & 3)) a[k >> 2] = rk ? -dinv : t;
    }
    if (bad) {
        if (tid == 0) *fail = 1;
        __syncthreads();
        return 1;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int c = w + 4 * s;
        if (i < R && c < R) Vinv[i + c * R] = -a[s];
    }
    __syncthreads();
    return 0;
}

}  // namespace tdx
