// vinv.hpp -- the per-mode inverse of a CP-ALS sweep, shared by the sweep driver (k_vinv)
// and the fused pass over T of the order-3 shared pass (k_m2_vinv): V = hadamard of the other
// modes' Grams, V^{-1} by the FP64 symmetric sweep operator, one workgroup of 256 threads.
#pragma once
#include "linalg.hpp"

namespace tdx {

static constexpr int kMaxR = 64;  // rank limit of the LDS-resident solve and of the inverse

// p[0] + p[stride] + ... + p[(n - 1) stride], added in index order, 16 loads in flight (the
// ones past n re-read a valid entry and are not added)
__device__ __forceinline__ double ordered_sum(const double *p, long stride, int n) {
    double s = 0.0;
    for (int b = 0; b < n; b += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = p[(long)(b + u < n ? b + u : n - 1) * stride];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (b + u < n) s += v[u];
    }
    return s;
}

// V = hadamard_{m != mode} G_m; Vinv = V^{-1} by the symmetric sweep operator (FP64, the
// register-row form of sweep_inverse_rows); if a pivot fails, a pseudo-inverse from the
// Jacobi eigendecomposition (eigenvalues below 1e-12 * max dropped) and status |= 1.
// NS: register slots per lane, R <= 4 NS (4, 8 or 16 for R <= 16, 32, 64).
// About 4 KB of LDS, so that the workgroup fits on a CU beside an MTTKRP or TTM workgroup
// (up to 136 KB): V is read from the Grams, and the rare fallback works in fb_ws (global,
// 2 R^2 doubles: a copy of V and the eigenvectors).
// V[i][c] = product of the other modes' Grams in mode order, ((G_a G_b) G_c), the value of
// multiplying them into 1.0 one mode after another.  A fill of a lane's register row issues
// every load first: load(s, i, c, in) loads slot s's entries (from entry (0, 0) when !in, so
// all addresses are valid), finish() waits once and combines (0 where !in).
// (A loop over the modes with a branch per mode, or the scheduler's default interleaving,
// waited for each load in turn: 32 round trips per lane at R = 64.)
struct VGather {
    const double *Ga, *Gb, *Gc;  // Gc is read and not used at order 3
    int R;
    bool four;
    double x[3][16];
    bool on[16];
    __device__ double operator()(int i, int c) const {  // one entry, loads in order
        const long e = i + (long)c * R;
        const double v = Ga[e] * Gb[e];
        return four ? v * Gc[e] : v;
    }
    // slot s (a compile-time index after unrolling) of the lane's row
    __device__ void load(int s, int i, int c, bool in) {
        const long e = in ? i + (long)c * R : 0;
        x[0][s] = Ga[e];
        x[1][s] = Gb[e];
        x[2][s] = Gc[e];
        on[s] = in;
    }
    template <int NS>
    __device__ void finish(double (&a)[NS]) {
        __builtin_amdgcn_sched_barrier(0);  // every load above is issued before the first use
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double v = x[0][s] * x[1][s];
            a[s] = on[s] ? (four ? v * x[2][s] : v) : 0.0;
        }
    }
};

// R <= 16: the same sweep operator in one wave, without barriers.  Lane l holds A[i][c] for
// row i = l % 16 and columns c = l / 16 + 4 s (s = 0..3).  A stays symmetric under the sweep,
// so pivot k's row equals its column, which sits in slot k / 4 of the lanes with
// l / 16 = k % 4 and reaches every lane by ds_bpermute.  Returns 1 (uniform) on a failed pivot.
template <class VAt>
__device__ inline int sweep_inverse_wave16(VAt &vat, double *Vinv, int R) {
    const int l = threadIdx.x & 63, i = l & 15, cg = l >> 4;
    double a[4], a0[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int c = cg + 4 * s;
        vat.load(s, i, c, i < R && c < R);
    }
    vat.finish(a);
#pragma unroll
    for (int s = 0; s < 4; ++s) a0[s] = a[s];
    int bad = 0;  // uniform: every lane reads the same pivot
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < R && !bad) {
            const int src = 16 * (k & 3);
            const double v = a[k >> 2], v0 = a0[k >> 2];
            const double d = __shfl(v, src + k), dk = __shfl(v0, src + k);
            const double ti = __shfl(v, src + i);  // A[i][k]
            double wr[4];                           // A[k][cg + 4 s] = A[cg + 4 s][k]
#pragma unroll
            for (int s = 0; s < 4; ++s) wr[s] = __shfl(v, src + cg + 4 * s);
            if (!(d > 1e-13 * dk) || !isfinite(d)) {
                bad = 1;
            } else {
                const double dinv = 1.0 / d;
                const double t = ti * dinv;
                const bool rk = i == k;
                const double alpha = rk ? 0.0 : 1.0, beta = rk ? dinv : -t;
#pragma unroll
                for (int s = 0; s < 4; ++s) a[s] = fma(beta, wr[s], alpha * a[s]);
                if (cg == (k & 3)) a[k >> 2] = rk ? -dinv : t;
            }
        }
    }
    if (bad) return 1;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int c = cg + 4 * s;
        if (i < R && c < R) Vinv[i + c * R] = -a[s];
    }
    return 0;
}

// The inverse for one workgroup of 256 threads (the body of k_vinv).  G: the current Grams
// (G[mode] unused).
template <int NS>
__device__ __forceinline__ void vinv_block(const double *const *G, int order, int mode, int R, double *Vinv, int *status,
                           double *fb_ws) {
    __shared__ double Wb[128];
    __shared__ double dg[kMaxR];
    __shared__ double sh[256];
    __shared__ double cs[4 * (kMaxR / 2 + 1)];
    __shared__ int fail;
    // V[i][c] = product of the other modes' Grams in mode order, ((G_a G_b) G_c), the value
    // of multiplying them into 1.0 one mode after another.  Straight-line loads (the third is
    // read and not used at order 3), so a caller's unrolled fill issues every load before its
    // first wait (a loop over the modes with a branch per mode waited for each load in turn).
    const double *Ga = (mode == 0) ? G[1] : G[0];
    const double *Gb = (mode <= 1) ? G[2] : G[1];
    const double *Gc = (order > 3) ? ((mode <= 2) ? G[3] : G[2]) : Ga;
    const bool four = order > 3;
    VGather vat{Ga, Gb, Gc, R, four};
    int bad;
    if constexpr (NS == 4) {
        if (threadIdx.x < 64) {
            const int b = sweep_inverse_wave16(vat, Vinv, R);
            if (threadIdx.x == 0) fail = b;
        }
        __syncthreads();
        bad = fail;
    } else {
        // (sweep_inverse_rows writes the diagonal into dg and starts with a barrier)
        bad = sweep_inverse_rows<NS>(vat, dg, Wb, Vinv, R, &fail);
    }
    if (bad) {
        double *L = fb_ws, *Q = fb_ws + R * R;
        for (int e = threadIdx.x; e < R * R; e += blockDim.x) L[e] = vat(e % R, e / R);  // rare path
        __syncthreads();
        pinv_jacobi(L, Q, Vinv, R, cs, sh);
        if (threadIdx.x == 0) atomicOr(status, 1);
    }
}

}  // namespace tdx
