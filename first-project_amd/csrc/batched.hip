// batched.hip -- CP-ALS of thousands of small, uneven 3rd-order tensors in one persistent
// kernel per sweep (no reference counterpart: the reference decomposes one tensor per call,
// src/cp_als.cu:3).
//
// Each workgroup pulls tensors from a size-sorted work queue (one atomic counter, reset by
// a one-thread kernel at the start of every sweep so the sweep can be replayed as a graph)
// and performs the whole Gauss-Seidel sweep of its tensor locally:
//   for n in 0..2:  M = MTTKRP_n streamed slab by slab (X(:,:,k0:k0+kc) in LDS, coalesced;
//                   FP32 MFMA: mode 0 M += slab * KR rows, modes 1, 2 T = slab^T * A then a
//                   segmented sum over the slab's columns),
//                   V = hadamard of the other Grams, FP64 sweep-operator inverse (Jacobi
//                   pseudo-inverse fallback), U_n = M V^{-1}, G_n = U_n^T U_n (FP64)
//   relerr from the fast-fit identity, appended to the tensor's history.
// Every tensor element is read from HBM once per mode, exactly like the single-tensor path.
// Tunables (block size, slab capacity) are template parameters picked by the autotuner.
#include "common.hpp"
#include "linalg.hpp"
#include <algorithm>
#include <map>
#include <stdio.h>
#include <string>
#include <string.h>
#include <vector>

namespace tdx {

static constexpr int kBMaxDim = 128;
static constexpr int kBMaxR = 16;
static constexpr int kSlabCols = 128;  // columns per slab (per-column scratch Ys)

struct BDesc {
    long xoff;
    long uoff[3];
    int I, J, K, R;
    long toff;  // split design with T: X x_0 A^T of the tensor (J K x R floats) at toff
};

typedef float bf32x4 __attribute__((ext_vector_type(4)));

// sum over the workgroup: wave shuffles, then one value per wave through LDS (sh >= BS/64
// doubles); every thread returns the total
template <int BS>
__device__ inline double bsum(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int tid = threadIdx.x;
    __syncthreads();
    if ((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) r += sh[w];
    return r;
}

__device__ inline int sel3(int m, int a, int b, int c) { return m == 0 ? a : (m == 1 ? b : c); }
__device__ inline long sel3l(int m, long a, long b, long c) { return m == 0 ? a : (m == 1 ? b : c); }

// The solve-phase arrays live in the slab (idle between the MTTKRP streams): V, L, W, V^{-1}
// (R^2 doubles each), the Jacobi fallback's rotation table and the reduction scratch.
template <int BS>
struct SolveLds {
    static constexpr int kDoubles = 4 * kBMaxR * kBMaxR + BS + 4 * (kBMaxR / 2 + 1);
};

// Workgroups per CU the variant is built for (the LDS of slabs up to 6K floats allows two)
template <int BS, int SLAB>
struct BatchedOcc {
    static constexpr int kWgs = (SLAB <= 6144) ? 2 : 1;
};

template <int BS, int SLAB>
__global__ __launch_bounds__(BS, (BatchedOcc<BS, SLAB>::kWgs)) void k_cp_batched_sweep(const float *__restrict__ X, float *__restrict__ U,
                                                         const BDesc *__restrict__ desc, const int *__restrict__ sched,
                                                         int ntensors, int *__restrict__ queue,
                                                         const double *__restrict__ normx2,
                                                         double *__restrict__ fit_hist, int max_sweeps,
                                                         const int *__restrict__ sweep_ctr, int *__restrict__ status) {
    static_assert(SLAB / 2 >= SolveLds<BS>::kDoubles, "the solve arrays fit in the slab");
    __shared__ __attribute__((aligned(16))) float slab[SLAB];
    __shared__ float Us[3][kBMaxDim * kBMaxR];
    __shared__ float Ut[kBMaxDim * kBMaxR];  // A transposed, [i][r] with r < 16 (zero padded)
    __shared__ float Ms[kBMaxDim * kBMaxR];
    __shared__ float Ys[kSlabCols * kBMaxR > kBMaxDim * kBMaxR ? kSlabCols * kBMaxR : kBMaxDim * kBMaxR];
    __shared__ double Gs[3][kBMaxR * kBMaxR];
    __shared__ double red[BS / 64];
    __shared__ int fail, slot;
    double *const V = reinterpret_cast<double *>(slab);
    double *const L = V + kBMaxR * kBMaxR;
    double *const W = L + kBMaxR * kBMaxR;
    double *const Vi = W + kBMaxR * kBMaxR;
    double *const sh = Vi + kBMaxR * kBMaxR;   // BS doubles (Jacobi fallback)
    double *const cs = sh + BS;
    const int tid = threadIdx.x;
    const int sweep = *sweep_ctr;
    constexpr int NW = BS / 64;
    const int lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
    for (;;) {
        if (tid == 0) slot = atomicAdd(queue, 1);
        __syncthreads();
        const int sl = __builtin_amdgcn_readfirstlane(slot);
        if (sl >= ntensors) break;  // uniform: every wave exits once the queue is drained
        // the descriptor in scalar registers (no per-thread copy of it)
        const int t = __builtin_amdgcn_readfirstlane(sched[sl]);
        const BDesc *dp = desc + t;
        const int I = dp->I, J = dp->J, K = dp->K, R = dp->R;
        const long xoff = dp->xoff, uo0 = dp->uoff[0], uo1 = dp->uoff[1], uo2 = dp->uoff[2];
        const float *x = X + xoff;
        // factors to LDS, Grams in FP64
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int Im = sel3(m, I, J, K);
            const long uo = sel3l(m, uo0, uo1, uo2);
            for (int e = tid; e < Im * R; e += BS) Us[m][e] = U[uo + e];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int Im = sel3(m, I, J, K);
            for (int e = tid; e < R * R; e += BS) {
                const int a = e % R, b = e / R;
                double s = 0.0;
                for (int i = 0; i < Im; ++i) s += (double)Us[m][i + a * Im] * (double)Us[m][i + b * Im];
                Gs[m][e] = s;
            }
        }
        __syncthreads();
        // slabs are runs of consecutive columns c = j + k*J of the mode-1 unfolding
        // (contiguous in memory), stored in LDS with column stride I + 1 (the MFMA operand
        // reads walk both indices); at most kSlabCols columns so the per-column scratch fits
        const int ncol = J * K;
        const int Ip = I + 1;
        const int CC = min(SLAB / Ip, kSlabCols);
        double ip_last = 0.0;
        for (int n = 0; n < 3; ++n) {
            const int In = sel3(n, I, J, K);
            const long uon = sel3l(n, uo0, uo1, uo2);
            for (int e = tid; e < In * R; e += BS) Ms[e] = 0.0f;
            if (n == 1) {
                // modes 1 and 2 contract i with the (just updated) A: A^T rows, zero padded
                for (int e = tid; e < kBMaxDim * kBMaxR; e += BS) {
                    const int i = e >> 4, r = e & 15;
                    Ut[e] = (i < I && r < R) ? Us[0][i + r * I] : 0.0f;
                }
            }
            // mode 0 accumulators: row tiles wv + NW u of M (16 rows x 16 columns each)
            bf32x4 acc0[(8 + NW - 1) / NW];
#pragma unroll
            for (int u = 0; u < (8 + NW - 1) / NW; ++u) acc0[u] = bf32x4{0.f, 0.f, 0.f, 0.f};
            // -------- MTTKRP, slab by slab, two slabs in flight: each slab's elements are
            // loaded into one of two register buffers two slabs ahead of their use --------
            // element e = tid + BS u of a slab is (i, q) = (e mod I, e div I); (i0, q0) for
            // u = 0 and the per-u step (di, dq) are computed once
            const int q0 = tid / I, i0 = tid - q0 * I, dq = BS / I, di = BS - dq * I;
            constexpr int EPT = (SLAB + BS - 1) / BS;
            float preA[EPT], preB[EPT];
            // Every fetch issues its loads unconditionally (a load issued on some paths only makes
            // the compiler's wait counting drain the queue): past the last slab every lane reads
            // the tensor's first element (one cached line).
            auto fetch = [&](float (&pre)[EPT], int cs0) {  // loads of the slab at column cs0
                const bool past = cs0 >= ncol;
                const int ccs = min(CC, ncol - cs0);
                const float *xs = x + (past ? 0L : (long)cs0 * I);  // uniform base, 32-bit offsets
                const int last = past ? 0 : ccs * I - 1;
#pragma unroll
                for (int u = 0; u < EPT; ++u) {
                    const int e = tid + BS * u;
                    pre[u] = xs[e < last ? e : last];
                }
            };
            auto step = [&](float (&pre)[EPT], int c0) {
                const int cc = min(CC, ncol - c0);
                {
                    int i = i0, q = q0;
#pragma unroll
                    for (int u = 0; u < EPT; ++u) {
                        if (q < cc) slab[i + q * Ip] = pre[u];
                        i += di;
                        q += dq;
                        if (i >= I) {
                            i -= I;
                            ++q;
                        }
                    }
                }
                if (n == 0) {
                    // Khatri-Rao rows of the slab's columns: Ys[q*16 + r] = B[j_q,r] C[k_q,r]
                    for (int e = tid; e < cc * kBMaxR; e += BS) {
                        const int q = e >> 4, r = e & 15, c = c0 + q;
                        Ys[e] = (r < R) ? Us[1][c % J + r * J] * Us[2][c / J + r * K] : 0.0f;
                    }
                }
                __syncthreads();
                fetch(pre, c0 + 2 * CC);
                if (n == 0) {
                    // M += slab (I x cc) * KR (cc x 16): A[i][q] = slab, B[q][r] = Ys
                    const int nks = (cc + 3) >> 2;
#pragma unroll
                    for (int u = 0; u < (8 + NW - 1) / NW; ++u) {
                        const int tt = wv + NW * u;
                        if (16 * tt < I) {
                            const int i = 16 * tt + lr;
                            for (int ks = 0; ks < nks; ++ks) {
                                const int q = 4 * ks + lk;
                                const float a = (i < I && q < cc) ? slab[i + q * Ip] : 0.0f;
                                const float b = (q < cc) ? Ys[q * kBMaxR + lr] : 0.0f;
                                acc0[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc0[u], 0, 0, 0);
                            }
                        }
                    }
                } else {
                    // T (cc x 16) = slab^T (cc x I) * A (I x 16) -> Ys
                    const int nks = (I + 3) >> 2;
                    for (int tt = wv; 16 * tt < cc; tt += NW) {
                        bf32x4 tacc = {0.f, 0.f, 0.f, 0.f};
                        const int q = 16 * tt + lr;
                        for (int ks = 0; ks < nks; ++ks) {
                            const int i = 4 * ks + lk;
                            const float a = (q < cc && i < I) ? slab[i + q * Ip] : 0.0f;
                            tacc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Ut[i * kBMaxR + lr], tacc, 0, 0, 0);
                        }
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int qq = 16 * tt + 4 * lk + v;
                            if (qq < cc) Ys[qq * kBMaxR + lr] = tacc[v];
                        }
                    }
                    __syncthreads();
                    if (n == 1) {
                        // M[j,r] += sum over the slab's columns with this j of T[q,r] C[k_q,r]
                        for (int e = tid; e < J * R; e += BS) {
                            const int j = e % J, r = e / J;
                            float a2 = 0.0f;
                            for (int q = (j - c0 % J + J) % J; q < cc; q += J)
                                a2 = fmaf(Ys[q * kBMaxR + r], Us[2][(c0 + q) / J + r * K], a2);
                            Ms[e] += a2;
                        }
                    } else {
                        // M[k,r] += sum over k's columns of T[q,r] B[j_q,r]: 16 lanes per (k, r),
                        // strided over the columns, then a shuffle reduction within the 16
                        const int kf = c0 / J, kl = (c0 + cc - 1) / J;
                        const int npair = (kl - kf + 1) * R;
                        const int g = tid & 15;
                        for (int p0 = 0; p0 < npair; p0 += BS / 16) {
                            const int pr = p0 + (tid >> 4);
                            float a2 = 0.0f;
                            int k = 0, r = 0;
                            if (pr < npair) {
                                k = kf + pr / R;
                                r = pr % R;
                                const int qa = max(0, k * J - c0), qb = min(cc, (k + 1) * J - c0);
                                for (int q = qa + g; q < qb; q += 16)
                                    a2 = fmaf(Ys[q * kBMaxR + r], Us[1][(c0 + q) % J + r * J], a2);
                            }
#pragma unroll
                            for (int o = 8; o > 0; o >>= 1) a2 += __shfl_xor(a2, o, 16);
                            if (pr < npair && g == 0) Ms[k + r * K] += a2;
                        }
                    }
                }
                __syncthreads();
            };
            fetch(preA, 0);
            fetch(preB, CC);
            for (int c0 = 0; c0 < ncol; c0 += 2 * CC) {
                step(preA, c0);
                step(preB, c0 + CC);  // past the last column: an empty step (no stores, no work)
            }
            if (n == 0) {
                // mode-0 accumulators -> M (row 4 (lane / 16) + v of the tile, column lane % 16)
#pragma unroll
                for (int u = 0; u < (8 + NW - 1) / NW; ++u) {
                    const int tt = wv + NW * u;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = 16 * tt + 4 * lk + v;
                        if (i < I && lr < R) Ms[i + lr * I] = acc0[u][v];
                    }
                }
                __syncthreads();
            }
            // -------- normal equations in FP64 (arrays in the idle slab) --------
            for (int e = tid; e < R * R; e += BS) {
                double v = 1.0;
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    if (m != n) v *= Gs[m][e];
                V[e] = v;
            }
            __syncthreads();
            if (block_spd_inverse<(kBMaxR * kBMaxR + BS - 1) / BS>(V, L, W, Vi, R, cs, sh, &fail) && tid == 0) atomicOr(status, 1);
            // U_n = M V^{-1}; <M, U_n> for the fit after the last mode
            double ip = 0.0;
            for (int e = tid; e < In * R; e += BS) {
                const int i = e % In, b = e / In;
                double s = 0.0;
                for (int a = 0; a < R; ++a) s += (double)Ms[i + a * In] * Vi[a + b * R];
                const float u = (float)s;
                Ys[e] = u;
                ip += (double)Ms[e] * (double)u;
            }
            __syncthreads();
            for (int e = tid; e < In * R; e += BS) {
                Us[n][e] = Ys[e];
                U[uon + e] = Ys[e];
            }
            if (n == 2) ip_last = bsum<BS>(ip, red);
            __syncthreads();
            for (int e = tid; e < R * R; e += BS) {
                const int a = e % R, b = e / R;
                double s = 0.0;
                for (int i = 0; i < In; ++i) s += (double)Us[n][i + a * In] * (double)Us[n][i + b * In];
                Gs[n][e] = s;
            }
            __syncthreads();
        }
        // fit: ||X - Xhat||^2 = ||X||^2 - 2 <M_2, U_2> + sum(G0 o G1 o G2)
        double hx = 0.0;
        for (int e = tid; e < R * R; e += BS) hx += Gs[0][e] * Gs[1][e] * Gs[2][e];
        hx = bsum<BS>(hx, red);
        if (tid == 0 && sweep < max_sweeps) {
            const double nx2 = normx2[t];
            fit_hist[(long)t * max_sweeps + sweep] = sqrt(fabs(nx2 - 2.0 * ip_last + hx) / nx2);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// The split design: one MTTKRP kernel and one solve kernel per mode for the whole batch.
// k_bm_mttkrp<MODE> carries only that mode's LDS (the slab, two factor blocks, the
// per-column scratch and, for modes 1 and 2, the accumulated M), so three or four
// workgroups fit on a CU where the persistent kernel's 78 KB allowed two; each workgroup
// walks the tensors of a queue and writes M (laid out like the factors) to HBM.
// k_bm_solve<MODE> (one workgroup per tensor) forms V from the per-tensor FP64 Grams,
// inverts it, writes U_n and G_n, and after mode 2 the fit.  The arithmetic, including every
// summation order, is the persistent kernel's with 256 threads: the results are the same bit
// for bit.

// G[t][m] = U_m^T U_m (FP64) for the first sweep (each solve keeps its G_n current)
__global__ __launch_bounds__(256) void k_bm_gram_init(const float *__restrict__ U, const BDesc *__restrict__ desc,
                                                      int n, double *__restrict__ G) {
    __shared__ float Us[kBMaxDim * kBMaxR];
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const BDesc d = desc[t];
        const int R = d.R;
        for (int m = 0; m < 3; ++m) {
            const int Im = sel3(m, d.I, d.J, d.K);
            for (int e = threadIdx.x; e < Im * R; e += 256) Us[e] = U[d.uoff[m] + e];
            __syncthreads();
            for (int e = threadIdx.x; e < R * R; e += 256) {
                const int a = e % R, b = e / R;
                double s = 0.0;
                for (int i = 0; i < Im; ++i) s += (double)Us[i + a * Im] * (double)Us[i + b * Im];
                G[((long)t * 3 + m) * (kBMaxR * kBMaxR) + e] = s;
            }
            __syncthreads();
        }
    }
}

// Tbuf != nullptr (mode 1): T = X x_0 A^T, which modes 1 and 2 share, is also written to HBM,
// one slab at a time, for k_bm_mode2_t
template <int MODE, int BS, int SLAB, int WGS>
__global__ __launch_bounds__(BS, WGS) void k_bm_mttkrp(const float *__restrict__ X, const float *__restrict__ U,
                                                       float *__restrict__ Mbuf, float *__restrict__ Tbuf,
                                                       const BDesc *__restrict__ desc,
                                                       const int *__restrict__ sched, int ntensors,
                                                       int *__restrict__ queue) {
    __shared__ __attribute__((aligned(16))) float slab[SLAB];
    __shared__ float Fa[kBMaxDim * kBMaxR];  // mode 0: B (J x R); modes 1, 2: A^T rows [i][16]
    __shared__ float Fb[kBMaxDim * kBMaxR];  // mode 0: C; mode 1: C; mode 2: B
    __shared__ float Ms[MODE == 0 ? 1 : kBMaxDim * kBMaxR];
    __shared__ float Ys[kSlabCols * kBMaxR];
    __shared__ int slot;
    const int tid = threadIdx.x;
    constexpr int NW = BS / 64;
    const int lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
    for (;;) {
        if (tid == 0) slot = atomicAdd(queue, 1);
        __syncthreads();
        const int sl = __builtin_amdgcn_readfirstlane(slot);
        if (sl >= ntensors) break;  // uniform
        const int t = __builtin_amdgcn_readfirstlane(sched[sl]);
        const BDesc *dp = desc + t;
        const int I = dp->I, J = dp->J, K = dp->K, R = dp->R;
        const long xoff = dp->xoff, uo0 = dp->uoff[0], uo1 = dp->uoff[1], uo2 = dp->uoff[2];
        const float *x = X + xoff;
        if constexpr (MODE == 0) {
            for (int e = tid; e < J * R; e += BS) Fa[e] = U[uo1 + e];
            for (int e = tid; e < K * R; e += BS) Fb[e] = U[uo2 + e];
        } else {
            for (int e = tid; e < kBMaxDim * kBMaxR; e += BS) {
                const int i = e >> 4, r = e & 15;
                Fa[e] = (i < I && r < R) ? U[uo0 + i + r * I] : 0.0f;
            }
            if constexpr (MODE == 1) {
                for (int e = tid; e < K * R; e += BS) Fb[e] = U[uo2 + e];
                for (int e = tid; e < J * R; e += BS) Ms[e] = 0.0f;
            } else {
                for (int e = tid; e < J * R; e += BS) Fb[e] = U[uo1 + e];
                for (int e = tid; e < K * R; e += BS) Ms[e] = 0.0f;
            }
        }
        __syncthreads();  // the factor blocks are read by every thread (mode 0: before the first step's barrier)
        const int ncol = J * K;
        const int Ip = I + 1;
        const int CC = min(SLAB / Ip, kSlabCols);
        bf32x4 acc0[(8 + NW - 1) / NW];
#pragma unroll
        for (int u = 0; u < (8 + NW - 1) / NW; ++u) acc0[u] = bf32x4{0.f, 0.f, 0.f, 0.f};
        const int q0 = tid / I, i0 = tid - q0 * I, dq = BS / I, di = BS - dq * I;
        constexpr int EPT = (SLAB + BS - 1) / BS;
        float preA[EPT], preB[EPT];
        auto fetch = [&](float (&pre)[EPT], int cs0) {
            const bool past = cs0 >= ncol;
            const int ccs = min(CC, ncol - cs0);
            const float *xs = x + (past ? 0L : (long)cs0 * I);
            const int last = past ? 0 : ccs * I - 1;
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                const int e = tid + BS * u;
                pre[u] = xs[e < last ? e : last];
            }
        };
        auto step = [&](float (&pre)[EPT], int c0) {
            const int cc = min(CC, ncol - c0);
            {
                int i = i0, q = q0;
#pragma unroll
                for (int u = 0; u < EPT; ++u) {
                    if (q < cc) slab[i + q * Ip] = pre[u];
                    i += di;
                    q += dq;
                    if (i >= I) {
                        i -= I;
                        ++q;
                    }
                }
            }
            if constexpr (MODE == 0) {
                for (int e = tid; e < cc * kBMaxR; e += BS) {
                    const int q = e >> 4, r = e & 15, c = c0 + q;
                    Ys[e] = (r < R) ? Fa[c % J + r * J] * Fb[c / J + r * K] : 0.0f;
                }
            }
            __syncthreads();
            fetch(pre, c0 + 2 * CC);
            if constexpr (MODE == 0) {
                const int nks = (cc + 3) >> 2;
#pragma unroll
                for (int u = 0; u < (8 + NW - 1) / NW; ++u) {
                    const int tt = wv + NW * u;
                    if (16 * tt < I) {
                        const int i = 16 * tt + lr;
                        for (int ks = 0; ks < nks; ++ks) {
                            const int q = 4 * ks + lk;
                            const float a = (i < I && q < cc) ? slab[i + q * Ip] : 0.0f;
                            const float b = (q < cc) ? Ys[q * kBMaxR + lr] : 0.0f;
                            acc0[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc0[u], 0, 0, 0);
                        }
                    }
                }
            } else {
                const int nks = (I + 3) >> 2;
                for (int tt = wv; 16 * tt < cc; tt += NW) {
                    bf32x4 tacc = {0.f, 0.f, 0.f, 0.f};
                    const int q = 16 * tt + lr;
                    for (int ks = 0; ks < nks; ++ks) {
                        const int i = 4 * ks + lk;
                        const float a = (q < cc && i < I) ? slab[i + q * Ip] : 0.0f;
                        tacc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Fa[i * kBMaxR + lr], tacc, 0, 0, 0);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int qq = 16 * tt + 4 * lk + v;
                        if (qq < cc) Ys[qq * kBMaxR + lr] = tacc[v];
                    }
                }
                __syncthreads();
                if constexpr (MODE == 1) {
                    if (Tbuf) {
                        float *tp = Tbuf + dp->toff + (long)c0 * R;
                        for (int e = tid; e < cc * R; e += BS) tp[e] = Ys[(e / R) * kBMaxR + e % R];
                    }
                    for (int e = tid; e < J * R; e += BS) {
                        const int j = e % J, r = e / J;
                        float a2 = 0.0f;
                        for (int q = (j - c0 % J + J) % J; q < cc; q += J)
                            a2 = fmaf(Ys[q * kBMaxR + r], Fb[(c0 + q) / J + r * K], a2);
                        Ms[e] += a2;
                    }
                } else {
                    const int kf = c0 / J, kl = (c0 + cc - 1) / J;
                    const int npair = (kl - kf + 1) * R;
                    const int g = tid & 15;
                    for (int p0 = 0; p0 < npair; p0 += BS / 16) {
                        const int pr = p0 + (tid >> 4);
                        float a2 = 0.0f;
                        int k = 0, r = 0;
                        if (pr < npair) {
                            k = kf + pr / R;
                            r = pr % R;
                            const int qa = max(0, k * J - c0), qb = min(cc, (k + 1) * J - c0);
                            for (int q = qa + g; q < qb; q += 16)
                                a2 = fmaf(Ys[q * kBMaxR + r], Fb[(c0 + q) % J + r * J], a2);
                        }
#pragma unroll
                        for (int o = 8; o > 0; o >>= 1) a2 += __shfl_xor(a2, o, 16);
                        if (pr < npair && g == 0) Ms[k + r * K] += a2;
                    }
                }
            }
            __syncthreads();
        };
        fetch(preA, 0);
        fetch(preB, CC);
        for (int c0 = 0; c0 < ncol; c0 += 2 * CC) {
            step(preA, c0);
            step(preB, c0 + CC);  // past the last column: an empty step
        }
        if constexpr (MODE == 0) {
#pragma unroll
            for (int u = 0; u < (8 + NW - 1) / NW; ++u) {
                const int tt = wv + NW * u;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 16 * tt + 4 * lk + v;
                    if (i < I && lr < R) Mbuf[uo0 + i + lr * I] = acc0[u][v];
                }
            }
        } else {
            const int In = MODE == 1 ? J : K;
            const long uon = MODE == 1 ? uo1 : uo2;
            for (int e = tid; e < In * R; e += BS) Mbuf[uon + e] = Ms[e];
        }
        __syncthreads();  // LDS free for the next tensor
    }
}

// Mode 2 from T (written by the mode-1 kernel): the slabs of T in the partition the mode-1
// kernel used (CC columns), each loaded into the per-column scratch and reduced with the same
// segmented sums, so M_2 equals the one streamed from X bit for bit.  The next slab's T is
// loaded into registers while the current one is reduced.
template <int BS, int SLAB, int WGS>
__global__ __launch_bounds__(BS, WGS) void k_bm_mode2_t(const float *__restrict__ X, const float *__restrict__ U,
                                                        float *__restrict__ Mbuf, float *__restrict__ Tbuf,
                                                        const BDesc *__restrict__ desc,
                                                        const int *__restrict__ sched, int ntensors,
                                                        int *__restrict__ queue) {
    (void)X;
    __shared__ float Fb[kBMaxDim * kBMaxR];  // B
    __shared__ float Ms[kBMaxDim * kBMaxR];
    __shared__ float Ys[kSlabCols * kBMaxR];
    __shared__ int slot;
    const int tid = threadIdx.x;
    constexpr int EPT = kSlabCols * kBMaxR / BS;
    for (;;) {
        if (tid == 0) slot = atomicAdd(queue, 1);
        __syncthreads();
        const int sl = __builtin_amdgcn_readfirstlane(slot);
        if (sl >= ntensors) break;  // uniform
        const int t = __builtin_amdgcn_readfirstlane(sched[sl]);
        const BDesc *dp = desc + t;
        const int I = dp->I, J = dp->J, K = dp->K, R = dp->R;
        const long uo1 = dp->uoff[1], uo2 = dp->uoff[2];
        const float *tb = Tbuf + dp->toff;
        for (int e = tid; e < J * R; e += BS) Fb[e] = U[uo1 + e];
        for (int e = tid; e < K * R; e += BS) Ms[e] = 0.0f;
        const int ncol = J * K;
        const int CC = min(SLAB / (I + 1), kSlabCols);
        // entry e = tid + BS u of a slab is (q, r) = (e / 16, e % 16) of the scratch; its T value
        // (0 past the slab or for r >= R) is loaded one slab ahead
        float pre[EPT];
        auto fetch = [&](int c0) {
            const int cc = min(CC, ncol - c0);
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                const int e = tid + BS * u, q = e >> 4, r = e & 15;
                pre[u] = (c0 < ncol && q < cc && r < R) ? tb[(long)(c0 + q) * R + r] : 0.0f;
            }
        };
        fetch(0);
        __syncthreads();  // B and the zeroed M
        for (int c0 = 0; c0 < ncol; c0 += CC) {
            const int cc = min(CC, ncol - c0);
#pragma unroll
            for (int u = 0; u < EPT; ++u) Ys[tid + BS * u] = pre[u];
            __syncthreads();
            fetch(c0 + CC);
            const int kf = c0 / J, kl = (c0 + cc - 1) / J;
            const int npair = (kl - kf + 1) * R;
            const int g = tid & 15;
            for (int p0 = 0; p0 < npair; p0 += BS / 16) {
                const int pr = p0 + (tid >> 4);
                float a2 = 0.0f;
                int k = 0, r = 0;
                if (pr < npair) {
                    k = kf + pr / R;
                    r = pr % R;
                    const int qa = max(0, k * J - c0), qb = min(cc, (k + 1) * J - c0);
                    for (int q = qa + g; q < qb; q += 16)
                        a2 = fmaf(Ys[q * kBMaxR + r], Fb[(c0 + q) % J + r * J], a2);
                }
#pragma unroll
                for (int o = 8; o > 0; o >>= 1) a2 += __shfl_xor(a2, o, 16);
                if (pr < npair && g == 0) Ms[k + r * K] += a2;
            }
            __syncthreads();
        }
        for (int e = tid; e < K * R; e += BS) Mbuf[uo2 + e] = Ms[e];
        __syncthreads();  // LDS free for the next tensor
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_bm_solve(float *__restrict__ U, const float *__restrict__ Mbuf,
                                                  double *__restrict__ G, const BDesc *__restrict__ desc, int ntensors,
                                                  const double *__restrict__ normx2, double *__restrict__ fit_hist,
                                                  int max_sweeps, const int *__restrict__ sweep_ctr,
                                                  int *__restrict__ status) {
    constexpr int BS = 256;
    __shared__ float Ms[kBMaxDim * kBMaxR];
    __shared__ float Un[kBMaxDim * kBMaxR];
    __shared__ double V[kBMaxR * kBMaxR], L[kBMaxR * kBMaxR], W[kBMaxR * kBMaxR], Vi[kBMaxR * kBMaxR];
    __shared__ double sh[BS], cs[4 * (kBMaxR / 2 + 1)];
    __shared__ double red[BS / 64];
    __shared__ int fail;
    const int tid = threadIdx.x;
    const int sweep = *sweep_ctr;
    for (int t = blockIdx.x; t < ntensors; t += gridDim.x) {
        const BDesc d = desc[t];
        const int R = d.R;
        const int In = sel3(MODE, d.I, d.J, d.K);
        const long uon = d.uoff[MODE];
        double *Gt = G + (long)t * 3 * (kBMaxR * kBMaxR);
        for (int e = tid; e < In * R; e += BS) Ms[e] = Mbuf[uon + e];
        for (int e = tid; e < R * R; e += BS) {
            double v = 1.0;
#pragma unroll
            for (int m = 0; m < 3; ++m)
                if (m != MODE) v *= Gt[m * (kBMaxR * kBMaxR) + e];
            V[e] = v;
        }
        __syncthreads();
        if (block_spd_inverse<(kBMaxR * kBMaxR + BS - 1) / BS>(V, L, W, Vi, R, cs, sh, &fail) && tid == 0)
            atomicOr(status, 1);
        double ip = 0.0;
        for (int e = tid; e < In * R; e += BS) {
            const int i = e % In, b = e / In;
            double s = 0.0;
            for (int a = 0; a < R; ++a) s += (double)Ms[i + a * In] * Vi[a + b * R];
            const float u = (float)s;
            Un[e] = u;
            ip += (double)Ms[e] * (double)u;
        }
        __syncthreads();
        for (int e = tid; e < In * R; e += BS) U[uon + e] = Un[e];
        double ip_last = 0.0;
        if (MODE == 2) ip_last = bsum<BS>(ip, red);
        double hx = 0.0;  // (mode 2) sum of G_0 o G_1 o G_2 over this thread's entries
        for (int e = tid; e < R * R; e += BS) {
            const int a = e % R, b = e / R;
            double s = 0.0;
            for (int i = 0; i < In; ++i) s += (double)Un[i + a * In] * (double)Un[i + b * In];
            Gt[MODE * (kBMaxR * kBMaxR) + e] = s;
            if (MODE == 2) hx += Gt[e] * Gt[kBMaxR * kBMaxR + e] * s;
        }
        if (MODE == 2) {
            hx = bsum<BS>(hx, red);
            if (tid == 0 && sweep < max_sweeps) {
                const double nx2 = normx2[t];
                fit_hist[(long)t * max_sweeps + sweep] = sqrt(fabs(nx2 - 2.0 * ip_last + hx) / nx2);
            }
        }
        __syncthreads();
    }
}

__global__ void k_batched_prologue(int *queue) { *queue = 0; }
__global__ void k_bm_prologue(int *queues) {
    queues[0] = 0;
    queues[1] = 0;
    queues[2] = 0;
}
__global__ void k_batched_epilogue(int *sweep_ctr) { *sweep_ctr += 1; }

__global__ void k_batched_norm2(const float *__restrict__ X, const BDesc *__restrict__ desc, int n,
                                double *__restrict__ normx2) {
    __shared__ double sh[256];
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const BDesc d = desc[t];
        const long N = (long)d.I * d.J * d.K;
        double s = 0.0, c = 0.0;
        for (long e = threadIdx.x; e < N; e += 256) {
            const double v = X[d.xoff + e];
            const double y = v * v - c, tt = s + y;
            c = (tt - s) - y;
            s = tt;
        }
        s = block_sum(s, sh);
        if (threadIdx.x == 0) normx2[t] = s;
    }
}

// U_m of tensor t <- U(0,1) of Philox stream (seed + t, m)
__global__ void k_batched_init(float *__restrict__ U, const BDesc *__restrict__ desc, int n, uint64_t seed) {
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const BDesc d = desc[t];
        const int dims[3] = {d.I, d.J, d.K};
        for (int m = 0; m < 3; ++m)
            for (long e = threadIdx.x; e < (long)dims[m] * d.R; e += blockDim.x)
                U[d.uoff[m] + e] = word_to_uniform(philox_word(seed + (uint64_t)t, (uint32_t)m, (uint64_t)e));
    }
}

// ---------------------------------------------------------------------------
typedef void (*BatchedKernel)(const float *, float *, const BDesc *, const int *, int, int *, const double *,
                              double *, int, const int *, int *);
typedef void (*BmMttkrp)(const float *, const float *, float *, float *, const BDesc *, const int *, int, int *);

struct Variant {
    int bs, slab;
    BatchedKernel fn;  // persistent design; nullptr: the split design (m[])
    const char *name;
    BmMttkrp m[3];
};

// buffers of the split design: M (laid out like the factors) and the per-tensor Grams
struct BmBuffers {
    float *M = nullptr;
    double *G = nullptr;
    float *T = nullptr;     // variants whose mode 2 reads T
    int *queues = nullptr;  // 3 work-queue counters
};

// LDS per workgroup: the slab plus ~54 KB of factors, Grams and per-mode scratch; the 4K-6K
// slab variants fit two workgroups per CU (160 KB of LDS), the larger ones one.
static const Variant kVariants[] = {
    {256, 4096, k_cp_batched_sweep<256, 4096>, "bs256_slab4k"},
    {256, 6144, k_cp_batched_sweep<256, 6144>, "bs256_slab6k"},
    {512, 4096, k_cp_batched_sweep<512, 4096>, "bs512_slab4k"},
    {512, 6144, k_cp_batched_sweep<512, 6144>, "bs512_slab6k"},
    {256, 8192, k_cp_batched_sweep<256, 8192>, "bs256_slab8k"},
    {512, 8192, k_cp_batched_sweep<512, 8192>, "bs512_slab8k"},
    {512, 12288, k_cp_batched_sweep<512, 12288>, "bs512_slab12k"},
    // split design: LDS = slab + 32 KB (mode 0: 24 KB), three workgroups per CU at 4K floats
    // (a 2K-float slab, four per CU, measured 4.2 ms on C4 against 3.3)
    {256, 4096, nullptr, "split_bs256_slab4k",
     {k_bm_mttkrp<0, 256, 4096, 3>, k_bm_mttkrp<1, 256, 4096, 3>, k_bm_mttkrp<2, 256, 4096, 3>}},
    {256, 6144, nullptr, "split_bs256_slab6k",
     {k_bm_mttkrp<0, 256, 6144, 2>, k_bm_mttkrp<1, 256, 6144, 2>, k_bm_mttkrp<2, 256, 6144, 2>}},
    // mode 2 from T = X x_0 A^T written by mode 1 (J K R floats per tensor, read instead of X)
    {256, 4096, nullptr, "split_t_bs256_slab4k",
     {k_bm_mttkrp<0, 256, 4096, 3>, k_bm_mttkrp<1, 256, 4096, 3>, k_bm_mode2_t<256, 4096, 3>}},
    {256, 6144, nullptr, "split_t_bs256_slab6k",
     {k_bm_mttkrp<0, 256, 6144, 2>, k_bm_mttkrp<1, 256, 6144, 2>, k_bm_mode2_t<256, 6144, 2>}},
};
static bool variant_uses_t(int v) { return v >= 0 && kVariants[v].m[2] != nullptr && strncmp(kVariants[v].name, "split_t_", 8) == 0; }
static constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

static int grid_for_kernel(const void *fn, int bs) {
    int per_cu = 2;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, bs, 0);
    if (per_cu < 1) per_cu = 1;
    return device_cus() * per_cu;
}

// one sweep of variant v (captured into the sweep graph)
static int launch_variant(int v, hipStream_t s, const float *X, float *U, const BDesc *desc, const int *sched, int n,
                          int *queue, const double *normx2, double *fit, int max_sweeps, int *sweep_ctr, int *status,
                          const BmBuffers &bm) {
    const Variant &V = kVariants[v];
    if (V.fn) {
        k_batched_prologue<<<1, 1, 0, s>>>(queue);
        hipLaunchKernelGGL(V.fn, dim3(grid_for_kernel((const void *)V.fn, V.bs)), dim3(V.bs), 0, s, X, U, desc, sched,
                           n, queue, normx2, fit, max_sweeps, sweep_ctr, status);
    } else {
        k_bm_prologue<<<1, 1, 0, s>>>(bm.queues);
        const int sgrid = n < 8192 ? n : 8192;
        for (int m = 0; m < 3; ++m) {
            hipLaunchKernelGGL(V.m[m], dim3(grid_for_kernel((const void *)V.m[m], V.bs)), dim3(V.bs), 0, s, X,
                               (const float *)U, bm.M, m >= 1 ? bm.T : nullptr, desc, sched, n, bm.queues + m);
            if (m == 0) k_bm_solve<0><<<sgrid, 256, 0, s>>>(U, bm.M, bm.G, desc, n, normx2, fit, max_sweeps, sweep_ctr, status);
            else if (m == 1) k_bm_solve<1><<<sgrid, 256, 0, s>>>(U, bm.M, bm.G, desc, n, normx2, fit, max_sweeps, sweep_ctr, status);
            else k_bm_solve<2><<<sgrid, 256, 0, s>>>(U, bm.M, bm.G, desc, n, normx2, fit, max_sweeps, sweep_ctr, status);
        }
    }
    k_batched_epilogue<<<1, 1, 0, s>>>(sweep_ctr);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

}  // namespace tdx

using namespace tdx;

extern "C" {

int tdx_cp_batched_variants(void) { return kNumVariants; }

const char *tdx_cp_batched_variant_name(int v) {
    return (v >= 0 && v < kNumVariants) ? kVariants[v].name : "";
}

// CP-ALS of `ntensors` 3rd-order tensors packed in X (tensor t at xoff[t], column-major
// dims[3t..3t+2]) with ranks[t] <= 16 and extents <= 128; factors packed in U (mode m of
// tensor t at uoff[3t+m], I_m x R column-major).  init = 0 draws U(0,1) from Philox stream
// (seed + t, m); 1 uses the caller's U.  variant < 0 picks variant 9 (split_t_bs256_slab4k,
// the autotuner's choice on the 4096-tensor benchmark batch).  fit_hist (host, may
// be NULL) receives ntensors x sweeps relative errors; sweep_ms (host, may be NULL) the
// device time of every sweep.  The per-sweep launches are captured once in a hipGraph.
int tdx_cp_als_batched(const float *X, float *U, const long *dims, const long *ranks, const long *xoff,
                       const long *uoff, long ntensors, int sweeps, int init, uint64_t seed, int variant,
                       double *fit_hist, float *sweep_ms) {
    if (!X || !U || !dims || !ranks || !xoff || !uoff || ntensors <= 0 || sweeps < 0) return record(TDX_ERR_ARG);
    if (variant >= kNumVariants) return record(TDX_ERR_ARG);
    if (variant < 0) variant = 9;
    std::vector<BDesc> h(ntensors);
    std::vector<std::pair<long, int>> order(ntensors);
    for (long t = 0; t < ntensors; ++t) {
        BDesc &d = h[t];
        d.I = (int)dims[3 * t];
        d.J = (int)dims[3 * t + 1];
        d.K = (int)dims[3 * t + 2];
        d.R = (int)ranks[t];
        if (d.I <= 0 || d.J <= 0 || d.K <= 0 || d.I > kBMaxDim || d.J > kBMaxDim || d.K > kBMaxDim || d.R <= 0 ||
            d.R > kBMaxR)
            return record(TDX_ERR_UNSUPPORTED);
        d.xoff = xoff[t];
        d.toff = 0;
        for (int m = 0; m < 3; ++m) d.uoff[m] = uoff[3 * t + m];
        order[t] = {-(long)d.I * d.J * d.K * d.R, (int)t};  // largest work first
    }
    std::sort(order.begin(), order.end());
    std::vector<int> sched(ntensors);
    for (long t = 0; t < ntensors; ++t) sched[t] = order[t].second;

    hipStream_t s = stream();
    hipStream_t owned = nullptr;
    if (!s) {
        TDX_HIP(hipStreamCreate(&owned));
        s = owned;
    }
    BDesc *d_desc = nullptr;
    int *d_sched = nullptr, *d_int = nullptr;
    double *d_nx = nullptr, *d_fit = nullptr;
    struct Guard {
        hipStream_t st;
        void *p[8];
        ~Guard() {
            if (st) (void)hipStreamSynchronize(st);
            for (void *q : p)
                if (q) (void)hipFree(q);
            if (st) (void)hipStreamDestroy(st);
        }
    } guard{owned, {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}};
    TDX_HIP(hipMalloc(&d_desc, sizeof(BDesc) * ntensors));
    guard.p[0] = d_desc;
    TDX_HIP(hipMalloc(&d_sched, sizeof(int) * ntensors));
    guard.p[1] = d_sched;
    TDX_HIP(hipMalloc(&d_int, sizeof(int) * 8));
    guard.p[2] = d_int;
    TDX_HIP(hipMalloc(&d_nx, sizeof(double) * ntensors));
    guard.p[3] = d_nx;
    const int S = sweeps > 0 ? sweeps : 1;
    TDX_HIP(hipMalloc(&d_fit, sizeof(double) * ntensors * S));
    guard.p[4] = d_fit;
    TDX_HIP(hipMemcpyAsync(d_desc, h.data(), sizeof(BDesc) * ntensors, hipMemcpyHostToDevice, s));
    TDX_HIP(hipMemcpyAsync(d_sched, sched.data(), sizeof(int) * ntensors, hipMemcpyHostToDevice, s));
    int *queue = d_int, *sweep_ctr = d_int + 1, *status = d_int + 2;
    k_batched_prologue<<<1, 1, 0, s>>>(sweep_ctr);
    k_batched_prologue<<<1, 1, 0, s>>>(status);
    if (init == 0) k_batched_init<<<1024, 256, 0, s>>>(U, d_desc, (int)ntensors, seed);
    k_batched_norm2<<<1024, 256, 0, s>>>(X, d_desc, (int)ntensors, d_nx);
    TDX_LAUNCH_CHECK();
    BmBuffers bm;
    if (!kVariants[variant].fn) {
        long mf = 0;  // M is laid out like the factors
        for (long t = 0; t < ntensors; ++t)
            for (int m = 0; m < 3; ++m) mf = std::max(mf, h[t].uoff[m] + (long)dims[3 * t + m] * h[t].R);
        TDX_HIP(hipMalloc(&bm.M, sizeof(float) * mf));
        guard.p[5] = bm.M;
        TDX_HIP(hipMalloc(&bm.G, sizeof(double) * ntensors * 3 * kBMaxR * kBMaxR));
        guard.p[6] = bm.G;
        if (variant_uses_t(variant)) {
            long tf = 0;
            for (long t = 0; t < ntensors; ++t) {
                h[t].toff = tf;
                tf += (long)h[t].J * h[t].K * h[t].R;
            }
            TDX_HIP(hipMalloc(&bm.T, sizeof(float) * tf));
            guard.p[7] = bm.T;
            TDX_HIP(hipMemcpyAsync(d_desc, h.data(), sizeof(BDesc) * ntensors, hipMemcpyHostToDevice, s));
        }
        bm.queues = d_int + 3;
        k_bm_gram_init<<<1024, 256, 0, s>>>(U, d_desc, (int)ntensors, bm.G);
        TDX_LAUNCH_CHECK();
    }
    // capture one sweep
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    TDX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = launch_variant(variant, s, X, U, d_desc, d_sched, (int)ntensors, queue, d_nx, d_fit, S, sweep_ctr,
                            status, bm);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return record_hip(e, "hipStreamEndCapture", __FILE__, __LINE__);
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return record_hip(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    std::vector<hipEvent_t> evs(sweeps + 1);
    for (auto &ev : evs) (void)hipEventCreate(&ev);
    (void)hipEventRecord(evs[0], s);
    for (int it = 0; it < sweeps && e == hipSuccess; ++it) {
        e = hipGraphLaunch(ge, s);
        if (e == hipSuccess) e = hipEventRecord(evs[it + 1], s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && sweep_ms)
        for (int it = 0; it < sweeps; ++it) (void)hipEventElapsedTime(&sweep_ms[it], evs[it], evs[it + 1]);
    for (auto &ev : evs) (void)hipEventDestroy(ev);
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return record_hip(e, "batched sweeps", __FILE__, __LINE__);
    if (fit_hist && sweeps > 0)
        TDX_HIP(hipMemcpy(fit_hist, d_fit, sizeof(double) * ntensors * S, hipMemcpyDeviceToHost));
    int hs = 0;
    TDX_HIP(hipMemcpy(&hs, status, sizeof(int), hipMemcpyDeviceToHost));
    return record(hs ? TDX_WARN_SINGULAR : TDX_OK);
}

// Autotuner: times `trial_sweeps` sweeps of every variant on copies of the initial factors
// (the caller's U is left as given) and returns the fastest variant in *best; the timings
// of all variants go to ms_per_sweep (host, tdx_cp_batched_variants() entries).  Results
// are cached in `cache_path` (a text file of "key variant ms" lines; NULL: no cache) under a
// key built from the batch shape and the device name.
int tdx_cp_batched_autotune(const float *X, const float *U, const long *dims, const long *ranks, const long *xoff,
                            const long *uoff, long ntensors, long u_floats, int trial_sweeps, const char *cache_path,
                            int *best, float *ms_per_sweep) {
    if (!X || !U || !best || ntensors <= 0 || u_floats <= 0 || trial_sweeps <= 0) return record(TDX_ERR_ARG);
    // cache key: device name + tensor count + total elements + the full rank histogram (1..16)
    // + an FNV-1a digest of the (extents, rank) list in order and of the variant names
    hipDeviceProp_t prop;
    int dev = 0;
    TDX_HIP(hipGetDevice(&dev));
    TDX_HIP(hipGetDeviceProperties(&prop, dev));
    long tot = 0, rh[17] = {0};
    uint64_t digest = 1469598103934665603ull;
    auto mix = [&](long v) {
        for (int b = 0; b < 8; ++b) {
            digest ^= (uint64_t)((v >> (8 * b)) & 0xff);
            digest *= 1099511628211ull;
        }
    };
    for (long t = 0; t < ntensors; ++t) {
        tot += dims[3 * t] * dims[3 * t + 1] * dims[3 * t + 2];
        if (ranks[t] >= 1 && ranks[t] <= 16) rh[ranks[t]]++;
        mix(dims[3 * t]);
        mix(dims[3 * t + 1]);
        mix(dims[3 * t + 2]);
        mix(ranks[t]);
    }
    // (the variant set is part of the key: a cache written by a build with other variants
    // never maps to a wrong index)
    for (int v = 0; v < kNumVariants; ++v)
        for (const char *c = kVariants[v].name; *c; ++c) mix((long)*c);
    char key[512];
    int kl = snprintf(key, sizeof(key), "%s|n%ld|e%ld|r", prop.gcnArchName, ntensors, tot);
    for (int r = 1; r <= 16 && kl > 0 && kl < (int)sizeof(key) - 24; ++r)
        kl += snprintf(key + kl, sizeof(key) - kl, "%s%ld", r > 1 ? "," : "", rh[r]);
    if (kl > 0 && kl < (int)sizeof(key) - 24) snprintf(key + kl, sizeof(key) - kl, "|d%016llx", (unsigned long long)digest);
    if (cache_path) {
        FILE *f = fopen(cache_path, "r");
        if (f) {
            char line[640];
            while (fgets(line, sizeof(line), f)) {
                char k[512];
                int v = -1;
                float ms = 0;
                if (sscanf(line, "%511s %d %f", k, &v, &ms) == 3 && std::string(k) == key && v >= 0 &&
                    v < kNumVariants) {
                    *best = v;
                    fclose(f);
                    return record(TDX_OK);
                }
            }
            fclose(f);
        }
    }
    float *Uw = nullptr;
    TDX_HIP(hipMalloc(&Uw, sizeof(float) * u_floats));
    float best_ms = 1e30f;
    int bv = 1, rc = TDX_OK;
    std::vector<float> ms(trial_sweeps);
    for (int v = 0; v < kNumVariants && !rc; ++v) {
        rc = (hipMemcpy(Uw, U, sizeof(float) * u_floats, hipMemcpyDeviceToDevice) == hipSuccess) ? TDX_OK
                                                                                              : record(TDX_ERR_HIP);
        if (rc) break;
        rc = tdx_cp_als_batched(X, Uw, dims, ranks, xoff, uoff, ntensors, trial_sweeps, 1, 0, v, nullptr, ms.data());
        if (rc == TDX_WARN_SINGULAR) rc = TDX_OK;
        float m = 0;
        for (float x : ms) m += x;
        m /= (float)trial_sweeps;
        if (ms_per_sweep) ms_per_sweep[v] = m;
        if (m < best_ms) {
            best_ms = m;
            bv = v;
        }
    }
    (void)hipFree(Uw);
    if (rc) return rc;
    *best = bv;
    if (cache_path) {
        FILE *f = fopen(cache_path, "a");
        if (f) {
            fprintf(f, "%s %d %.5f\n", key, bv, best_ms);
            fclose(f);
        }
    }
    return record(TDX_OK);
}

}  // extern "C"
