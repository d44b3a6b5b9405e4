// tucker.hip -- Tucker HOSVD and HOOI for gfx950.
//
// HOSVD (src/tucker_hosvd.cu:3-81, src/tucker4_hosvd.cu:3-121): per mode the Gram of the
// unfolding X(n) X(n)^T, a symmetric eigensolve, and the eigenvectors of the r_n largest
// eigenvalues in ascending order (the reference copies the last r columns of the
// ascending syevdx output, src/tucker_hosvd.cu:68-70); core = X x_1 U_1^T ... x_d U_d^T.
// HOOI (decomposition/GPU/normal/tucker/tucker_hooi.cu:103-168, generalised to order 4):
// per mode, Y = X x_{m != n} U_m^T, then the same Gram + eigensolve on Y(n).
//
// The Gram never forms an unfolded copy (the reference materialises X2 with t2m and, for
// order 4, X2 and X3 by identity-matrix GEMMs, src/tucker4_hosvd.cu:23-37): a workgroup
// stages 32 fibres of the (P, In, Q) view in LDS and accumulates an In x In FP32 tile,
// partial tiles are summed in FP64 in a fixed order.  The eigensolve is the FP64
// tridiagonal solver in one workgroup (subspace iteration first where the top eigenvalues
// stand clear of the rest, warm-started from the previous factor in HOOI); eigenvector signs
// are fixed so that the entry of largest magnitude is positive (syevdx leaves the sign
// unspecified).
#include "common.hpp"
#include "eig.hpp"
#include "eig_sub.hpp"
#include "linalg.hpp"
#include <vector>

namespace tdx {

int ttm_generic(const float *X, int order, const long *dims, int mode, const float *U, long lda, long ldi, long Ra,
                float *Y, hipStream_t s);
int gram_mfma_launch(const void *X, int xtype, long P, long In, long Q, float *part, int max_parts, hipStream_t s);
int ttm_generic_t(const void *X, int xtype, int order, const long *dims, int mode, const float *U, long lda, long ldi,
                  long Ra, float *Y, hipStream_t s);

static constexpr int kGramMaxN = 128;
static constexpr int kGramChunk = 32;

// partial Gram of the (P, In, Q) view over fibre chunks: thread (ty, tx) owns an 8x8 block
// of the In x In tile (In <= 128).  Fibre k of a chunk is (p, q) with p fastest.
__device__ __forceinline__ float ldx(const float *p, long i) { return p[i]; }
__device__ __forceinline__ float ldx(const uint16_t *p, long i) { return __uint_as_float((uint32_t)p[i] << 16); }

template <typename T>
__global__ __launch_bounds__(256) void k_gram_partial(const T *__restrict__ X, long P, long In, long Q,
                                                      long kper, float *__restrict__ part) {
    __shared__ float xs[kGramChunk][kGramMaxN + 1];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const long K = P * Q;
    const long k0 = blockIdx.x * kper, k1 = min(k0 + kper, K);
    float acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0f;
    for (long kb = k0; kb < k1; kb += kGramChunk) {
        // stage X[fibre kb+kk][a] for kk < 32, a < In
        for (int e = tid; e < kGramChunk * kGramMaxN; e += 256) {
            int kk, a;
            if (P == 1) { a = e % kGramMaxN; kk = e / kGramMaxN; }   // contiguous along a
            else { kk = e % kGramChunk; a = e / kGramChunk; }      // contiguous along p
            const long k = kb + kk;
            float v = 0.0f;
            if (k < k1 && a < In) {
                const long p = k % P, q = k / P;
                v = ldx(X, p + a * P + q * P * In);
            }
            xs[kk][a] = v;
        }
        __syncthreads();
        for (int kk = 0; kk < kGramChunk; ++kk) {
            float av[8], bv[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                av[i] = xs[kk][ty * 8 + i];
                bv[i] = xs[kk][tx * 8 + i];
            }
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] = fmaf(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
    float *dst = part + (long)blockIdx.x * kGramMaxN * kGramMaxN;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) dst[(ty * 8 + i) * kGramMaxN + tx * 8 + j] = acc[i][j];
}

// G = the sum of the nblk FP32 partial Grams (pitch kGramMaxN), in FP64: a workgroup takes 64
// consecutive entries (a, b) of the upper triangle (b fastest: contiguous reads) and its four
// waves the four quarters of the block range (each summed in block order, 8 loads in
// flight); the quarters are added in order and the sum written to both triangles, so G is
// symmetric bit for bit.  Grid: ceil(In^2 / 64).
__global__ __launch_bounds__(256) void k_gram_sum(const float *__restrict__ part, int nblk, long In,
                                                  double *__restrict__ G) {
    __shared__ double sh[4][64];
    const int lane = threadIdx.x & 63, rg = threadIdx.x >> 6;
    const long f = blockIdx.x * 64L + lane;
    const long a = f / In, b = f - a * In;
    const bool on = f < In * In && a <= b;
    double s = 0.0;
    if (on) {
        const int k0 = (int)((long)rg * nblk / 4), k1 = (int)((long)(rg + 1) * nblk / 4);
        const float *p = part + a * kGramMaxN + b;
        constexpr long kStride = (long)kGramMaxN * kGramMaxN;
        for (int k = k0; k < k1; k += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = p[(long)(k + u < k1 ? k + u : k1 - 1) * kStride];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + u < k1) s += (double)v[u];
        }
    }
    sh[rg][lane] = s;
    __syncthreads();
    if (rg == 0 && on) {
        const double t = ((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane];
        G[a + b * In] = t;
        G[b + a * In] = t;
    }
}

// Top-r eigenvectors of the symmetric n x n G (FP64), one workgroup: U (n x r, col-major
// FP32) receives the eigenvectors of the r largest eigenvalues in ascending order.  With
// try_sub the subspace-iteration path runs first (start block U0, or a built-in one when U0
// is null; U0 may alias U) and the tridiagonal solver only when it declines.  method (optional,
// device) receives 1 for the subspace path, 0 for the tridiagonal one.
__global__ __launch_bounds__(kEigThreads) void k_top_eigvecs(const double *__restrict__ G, int n, int r,
                                                             const float *U0, int try_sub, float *U, double *w,
                                                             double *__restrict__ scratch, int *method, int rotate) {
    extern __shared__ double smem[];
    if (try_sub && r <= kSubMaxR) {
        if (sym_top_eigvecs_subspace(G, n, r, U0, U, w, smem, rotate != 0)) {
            if (method && threadIdx.x == 0) *method = 1;
            return;
        }
        __syncthreads();  // the LDS is reused below
    }
    sym_top_eigvecs_block(G, n, r, U, w, scratch, scratch + (long)n * n, smem);
    if (method && threadIdx.x == 0) *method = 0;
}

__global__ void k_sumsq(const float *__restrict__ X, long n, double *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long e = threadIdx.x; e < n; e += 256) s += (double)X[e] * (double)X[e];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_norm2_blocks(const T *__restrict__ X, long n, double *__restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0, c = 0.0;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        double x = (double)X[e];
        double y = x * x - c, t = s + y;
        c = (t - s) - y;
        s = t;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_sum_to(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = s;
}

// ---------------------------------------------------------------------------
struct TuckerWs {
    float *gram_part = nullptr;
    int gram_blocks = 0;
    double *G = nullptr;
    double *Q = nullptr;
    float *y0 = nullptr, *y1 = nullptr, *y2 = nullptr;
    double *scal = nullptr;
    ~TuckerWs() {
        void *ps[] = {gram_part, G, Q, y0, y1, y2, scal};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
};

static int unfold_gram(const void *X, int xtype, int order, const long *dims, int mode, TuckerWs &W, double *G,
                       hipStream_t s) {
    long P = 1, Q = 1;
    for (int m = 0; m < mode; ++m) P *= dims[m];
    for (int m = mode + 1; m < order; ++m) Q *= dims[m];
    const long In = dims[mode];
    if (In > kGramMaxN) return record(TDX_ERR_UNSUPPORTED);
    const long K = P * Q;
    // MFMA path for the shapes it covers (16-byte vectors along the contiguous index)
    {
        const int mb = W.gram_blocks < 512 ? W.gram_blocks : 512;
        const int nb = gram_mfma_launch(X, xtype, P, In, Q, W.gram_part, mb, s);
        if (nb < 0) return record(TDX_ERR_HIP);
        if (nb > 0) {
            k_gram_sum<<<(int)ceil_div(In * In, 64L), 256, 0, s>>>(W.gram_part, nb, In, G);
            TDX_LAUNCH_CHECK();
            return TDX_OK;
        }
    }
    long nblk = ceil_div(K, 2048);
    if (nblk > W.gram_blocks) nblk = W.gram_blocks;
    if (nblk < 1) nblk = 1;
    long kper = ceil_div(ceil_div(K, nblk), kGramChunk) * kGramChunk;
    nblk = ceil_div(K, kper);
    if (xtype == TDX_BF16)
        k_gram_partial<uint16_t><<<(int)nblk, 256, 0, s>>>((const uint16_t *)X, P, In, Q, kper, W.gram_part);
    else
        k_gram_partial<float><<<(int)nblk, 256, 0, s>>>((const float *)X, P, In, Q, kper, W.gram_part);
    k_gram_sum<<<(int)ceil_div(In * In, 64L), 256, 0, s>>>(W.gram_part, (int)nblk, In, G);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

// TDX_EIG_SUBSPACE=0 keeps every eigensolve on the tridiagonal path (A/B comparisons)
static bool subspace_enabled() {
    static int on = -1;
    if (on < 0) {
        const char *e = getenv("TDX_EIG_SUBSPACE");
        on = (e && e[0] == '0') ? 0 : 1;
    }
    return on != 0;
}

static long eig_kernel_lds_doubles(int n) {
    const long a = eig_lds_doubles(n), b = eig_sub_doubles(n);
    return a > b ? a : b;
}

// scratch: 2 n^2 doubles.  try_sub: try the subspace path first, starting from U0 (or the
// built-in block when U0 is null).  rotate = 0: an orthonormal basis of the top-r subspace is
// enough (the subspace path may then skip its Rayleigh-Ritz rotation; w is not written then).
static int top_eigvecs(const double *G, long n, long r, double *scratch, const float *U0, int try_sub, float *U,
                       double *w, int *method, hipStream_t s, int rotate = 1) {
    if (n < 1 || n > kEigMaxN || r < 1 || r > n) return record(TDX_ERR_ARG);
    static bool attr_set = false;
    if (!attr_set) {
        TDX_HIP(hipFuncSetAttribute((const void *)k_top_eigvecs, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(double) * eig_kernel_lds_doubles(kEigMaxN))));
        attr_set = true;
    }
    const int sub = (try_sub && subspace_enabled()) ? 1 : 0;
    k_top_eigvecs<<<1, kEigThreads, sizeof(double) * eig_kernel_lds_doubles((int)n), s>>>(
        G, (int)n, (int)r, U0, sub, U, w, scratch, method, rotate);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

static int ws_setup(TuckerWs &W, int order, const long *dims, const long *ranks) {
    W.gram_blocks = 1024;
    TDX_HIP(hipMalloc(&W.gram_part, sizeof(float) * (size_t)W.gram_blocks * kGramMaxN * kGramMaxN));
    TDX_HIP(hipMalloc(&W.G, sizeof(double) * kGramMaxN * kGramMaxN));
    TDX_HIP(hipMalloc(&W.Q, sizeof(double) * 2 * kGramMaxN * kGramMaxN));
    TDX_HIP(hipMalloc(&W.scal, sizeof(double) * 2048));
    // largest intermediate of a TTM chain: the first contraction of the biggest mode
    long N = 1;
    for (int m = 0; m < order; ++m) N *= dims[m];
    long maxsz = 0;
    for (int m = 0; m < order; ++m) {
        long sz = N / dims[m] * ranks[m];
        if (sz > maxsz) maxsz = sz;
    }
    TDX_HIP(hipMalloc(&W.y0, sizeof(float) * maxsz));
    TDX_HIP(hipMalloc(&W.y1, sizeof(float) * maxsz));
    TDX_HIP(hipMalloc(&W.y2, sizeof(float) * maxsz));
    return TDX_OK;
}

// Y = X x_{m in modes} U_m^T applied in the listed order, results alternating between the
// buffers a and b (X itself when nmodes == 0); cur: the dims of X in, of the result out
static int ttm_seq(const void *X, int xtype, int order, long *cur, const long *ranks, float *const *U,
                   const int *modes, int nmodes, float *a, float *b, const float **out, hipStream_t s) {
    const void *src = X;
    for (int t = 0; t < nmodes; ++t) {
        const int m = modes[t];
        float *dst = (t & 1) ? b : a;
        int st = ttm_generic_t(src, t == 0 ? xtype : TDX_F32, order, cur, m, U[m], cur[m], 1, ranks[m], dst, s);
        if (st) return st;
        cur[m] = ranks[m];
        src = dst;
    }
    *out = (const float *)src;
    return TDX_OK;
}

// core = X x_1 U_1^T ... x_d U_d^T from the FP32 tensor, slowest mode first
static int core_from(const float *X, int order, const long *dims, const long *ranks, float *const *U, TuckerWs &W,
                     float *core, hipStream_t s) {
    int modes[4], nm = 0;
    for (int m = order - 1; m >= 1; --m) modes[nm++] = m;
    long cur[4];
    for (int m = 0; m < order; ++m) cur[m] = dims[m];
    const float *Y = nullptr;
    int st = ttm_seq(X, TDX_F32, order, cur, ranks, U, modes, nm, W.y0, W.y1, &Y, s);
    if (st) return st;
    return ttm_generic(Y, order, cur, 0, U[0], dims[0], 1, ranks[0], core, s);
}

static int tucker_run(const float *X, int order, const long *dims, const long *ranks, int iters, float *core,
                      float *const *U, double *relerr_hist, int bf16) {
    if (!X || !core || !U || order < 2 || order > 4) return record(TDX_ERR_ARG);
    for (int m = 0; m < order; ++m) {
        if (dims[m] <= 0 || ranks[m] <= 0 || ranks[m] > dims[m] || !U[m]) return record(TDX_ERR_ARG);
        if (dims[m] > kGramMaxN) return record(TDX_ERR_UNSUPPORTED);
    }
    hipStream_t s = stream();
    TuckerWs W;
    int st = ws_setup(W, order, dims, ranks);
    if (st) return st;
    long N = 1;
    for (int m = 0; m < order; ++m) N *= dims[m];
    // BF16 storage: the Grams of X and the first TTM of every chain stream a BF16 copy
    const void *Xs = X;
    int xtype = TDX_F32;
    uint16_t *Xb = nullptr;
    struct FreeB {
        uint16_t *&p;
        ~FreeB() {
            if (p) (void)hipFree(p);
        }
    } freeb{Xb};
    if (bf16) {
        TDX_HIP(hipMalloc(&Xb, sizeof(uint16_t) * N));
        st = tdx_quantize_bf16(X, Xb, N);
        if (st) return st;
        Xs = Xb;
        xtype = TDX_BF16;
    }
    // HOSVD init (only the spans matter when HOOI iterations follow: every later product
    // contracts U U^T, so the eigenvector rotation is left to the last iteration)
    for (int m = 0; m < order; ++m) {
        st = unfold_gram(Xs, xtype, order, dims, m, W, W.G, s);
        if (!st) st = top_eigvecs(W.G, dims[m], ranks[m], W.Q, nullptr, 1, U[m], nullptr, nullptr, s, iters == 0);
        if (st) return st;
    }
    k_norm2_blocks<float><<<1024, 256, 0, s>>>(X, N, W.scal);
    k_sum_to<<<1, 256, 0, s>>>(W.scal, 1024, W.scal + 1024);
    TDX_LAUNCH_CHECK();
    double nx2 = 0.0;
    long nc = 1;
    for (int m = 0; m < order; ++m) nc *= ranks[m];
    // HOOI with a dimension tree: the modes split into L = {0..h-1} and R = {h..d-1};
    // T = X x_{R} U^T serves every n in L, then T = X x_{L} U^T (with the new U_L) every n in
    // R, so X is streamed twice per iteration rather than d times.  Each T contracts its
    // slowest mode first.
    const int h = order / 2;
    for (int it = 0; it < iters; ++it) {
        const float *Y = nullptr;
        long ydims[4];
        for (int half = 0; half < 2; ++half) {
            const int n0 = half ? h : 0, n1 = half ? order : h;
            int tm[4], nt = 0;
            for (int m = order - 1; m >= 0; --m)
                if (m < n0 || m >= n1) tm[nt++] = m;
            long tdims[4];
            for (int m = 0; m < order; ++m) tdims[m] = dims[m];
            const float *T = nullptr;
            st = ttm_seq(Xs, xtype, order, tdims, ranks, U, tm, nt, W.y0, W.y1, &T, s);
            if (st) return st;
            float *fa = (T == W.y0) ? W.y1 : W.y0;
            for (int n = n0; n < n1; ++n) {
                int sm[4], ns = 0;
                for (int m = n1 - 1; m >= n0; --m)
                    if (m != n) sm[ns++] = m;
                for (int m = 0; m < order; ++m) ydims[m] = tdims[m];
                st = ttm_seq(T, TDX_F32, order, ydims, ranks, U, sm, ns, W.y2, fa, &Y, s);
                if (!st) st = unfold_gram(Y, TDX_F32, order, ydims, n, W, W.G, s);
                // warm start: the factor of the previous iteration spans nearly the same subspace;
                // eigenvectors (ascending, sign convention) only in the last iteration
                if (!st)
                    st = top_eigvecs(W.G, dims[n], ranks[n], W.Q, U[n], 1, U[n], nullptr, nullptr, s, it == iters - 1);
                if (st) return st;
            }
        }
        if (relerr_hist) {
            // relerr^2 = 1 - ||core||^2/||X||^2.  The core comes from the last Y (all modes
            // but d-1 contracted); with BF16 storage it is recomputed from the FP32 tensor.
            if (bf16) st = core_from(X, order, dims, ranks, U, W, core, s);
            else st = ttm_generic(Y, order, ydims, order - 1, U[order - 1], dims[order - 1], 1, ranks[order - 1],
                                  core, s);
            if (st) return st;
            {
                // ||core||^2: block partials (the first 1024 slots are free once ||X||^2 is summed)
                const int nb = (int)(ceil_div(nc, 1024L) < 1024 ? ceil_div(nc, 1024L) : 1024);
                k_norm2_blocks<float><<<nb, 256, 0, s>>>(core, nc, W.scal);
                k_sum_to<<<1, 256, 0, s>>>(W.scal, nb, W.scal + 1025);
            }
            double hv[2];
            TDX_HIP(hipMemcpyAsync(hv, W.scal + 1024, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
            TDX_HIP(hipStreamSynchronize(s));
            nx2 = hv[0];
            relerr_hist[it] = sqrt(fmax(nx2 - hv[1], 0.0) / nx2);
        }
    }
    // final core, always from the FP32 tensor
    st = core_from(X, order, dims, ranks, U, W, core, s);
    if (st) return st;
    TDX_HIP(hipStreamSynchronize(s));
    return TDX_OK;
}

}  // namespace tdx

using namespace tdx;

extern "C" {

int tdx_unfold_gram_ex(const void *X, int dtype, int order, const long *dims, int mode, double *G) {
    TDX_CHECK_ARG(X && dims && G && order >= 2 && order <= 4 && mode >= 0 && mode < order);
    TDX_CHECK_ARG(dtype == TDX_F32 || dtype == TDX_BF16);
    for (int m = 0; m < order; ++m) TDX_CHECK_ARG(dims[m] > 0);
    TuckerWs W;
    W.gram_blocks = 1024;
    TDX_HIP(hipMalloc(&W.gram_part, sizeof(float) * (size_t)W.gram_blocks * kGramMaxN * kGramMaxN));
    int st = unfold_gram(X, dtype, order, dims, mode, W, G, stream());
    if (!st) TDX_HIP(hipStreamSynchronize(stream()));
    return record(st);
}

int tdx_unfold_gram(const float *X, int order, const long *dims, int mode, double *G) {
    return tdx_unfold_gram_ex(X, TDX_F32, order, dims, mode, G);
}

int tdx_sym_eig_topr_ex(const double *G, long n, long r, const float *U0, int flags, float *U, double *w,
                        int *method) {
    TDX_CHECK_ARG(G && U && n >= 1 && n <= kEigMaxN && r >= 1 && r <= n);
    void *scr = nullptr;
    int st = workspace(WS_MISC2, sizeof(double) * 2 * n * n + 256, &scr);
    if (st) return st;
    int *dmethod = (int *)((char *)scr + sizeof(double) * 2 * n * n);
    st = top_eigvecs(G, n, r, (double *)scr, U0, (flags & TDX_EIG_TRY_SUBSPACE) ? 1 : 0, U, w, dmethod, stream());
    int hm = 0;
    if (!st) TDX_HIP(hipMemcpyAsync(&hm, dmethod, sizeof(int), hipMemcpyDeviceToHost, stream()));
    if (!st) TDX_HIP(hipStreamSynchronize(stream()));
    if (!st && method) *method = hm;
    return record(st);
}

int tdx_sym_eig_topr(const double *G, long n, long r, float *U, double *w) {
    return tdx_sym_eig_topr_ex(G, n, r, nullptr, 0, U, w, nullptr);
}

int tdx_tucker_hooi(const float *X, int order, const long *dims, const long *ranks, int iters, int bf16_ttm,
                    float *core, float *const *U, double *relerr_hist) {
    if (!dims || !ranks || iters < 0) return record(TDX_ERR_ARG);
    return record(tucker_run(X, order, dims, ranks, iters, core, U, relerr_hist, bf16_ttm));
}

int tdx_tucker_hosvd(const float *X, float *core, float *U1, float *U2, float *U3, long a, long b, long c, long r1,
                     long r2, long r3) {
    const long dims[3] = {a, b, c}, ranks[3] = {r1, r2, r3};
    float *U[3] = {U1, U2, U3};
    return record(tucker_run(X, 3, dims, ranks, 0, core, U, nullptr, 0));
}

int tdx_tucker_hosvd4(const float *X, float *core, float *U1, float *U2, float *U3, float *U4, long a, long b, long c,
                      long d, long r1, long r2, long r3, long r4) {
    const long dims[4] = {a, b, c, d}, ranks[4] = {r1, r2, r3, r4};
    float *U[4] = {U1, U2, U3, U4};
    return record(tucker_run(X, 4, dims, ranks, 0, core, U, nullptr, 0));
}

// reference tucker_tensorcore (src/tucker_tensorcore.cu): reduced-precision Grams and TTM
// chain; here the Grams and the first TTM stream a BF16 copy, everything accumulates in FP32
int tdx_tucker_tensorcore(const float *X, float *core, float *U1, float *U2, float *U3, long a, long b, long c,
                          long r1, long r2, long r3) {
    const long dims[3] = {a, b, c}, ranks[3] = {r1, r2, r3};
    float *U[3] = {U1, U2, U3};
    return record(tucker_run(X, 3, dims, ranks, 0, core, U, nullptr, 1));
}

}  // extern "C"
