// recon.hip -- reconstruction, residual norms, synthetic CP tensors and the generic
// mode-n product used by the Tucker path.
#include "common.hpp"
#include <math.h>

namespace tdx {

__device__ double block_sum_d(double v, double *sh) {
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

struct FacArgs {
    const float *U[4];
    long dims[4];
    int order;
    int R;
};

// For each fibre along mode 0 (fixed i_1..i_{d-1}): w[r] = prod_{m>=1} U_m[i_m, r] in
// LDS, then x_hat[i] = sum_r U_0[i, r] w[r].
//  MODE 0: write x_hat (gencp / gencp4, src/rse.cu:17-30,53-73)
//  MODE 1: accumulate (x - x_hat)^2 and x^2 in FP64 (compensated CP residual)
//  MODE 2: write x_hat and accumulate x_hat^2 and e^2 for E ~ N(0,1) (synthetic tensors)
template <int MODE>
__global__ __launch_bounds__(256) void k_cp_fibres(FacArgs f, float *__restrict__ out, const float *__restrict__ X,
                                                   double *__restrict__ part, uint64_t seed, uint32_t strm) {
    __shared__ float w[64 * 4];
    __shared__ double sh[256];
    const long I0 = f.dims[0];
    long nfib = 1;
    for (int m = 1; m < f.order; ++m) nfib *= f.dims[m];
    double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0;
    for (long fib = blockIdx.x; fib < nfib; fib += gridDim.x) {
        __syncthreads();
        for (int r = threadIdx.x; r < f.R; r += blockDim.x) {
            long rem = fib;
            float p = 1.0f;
            // reference order: the KR of the later modes is formed first (src/rse.cu:65-66)
            float q = 1.0f;
            for (int m = 1; m < f.order; ++m) {
                long idx = rem % f.dims[m];
                rem /= f.dims[m];
                if (m == 1) p = f.U[1][idx + r * f.dims[1]];
                else q = (m == 2) ? f.U[2][idx + r * f.dims[2]] : q * f.U[m][idx + r * f.dims[m]];
            }
            w[r] = (f.order > 2) ? p * q : p;
        }
        __syncthreads();
        for (long i = threadIdx.x; i < I0; i += blockDim.x) {
            float acc = 0.0f;
            for (int r = 0; r < f.R; ++r) acc = fmaf(f.U[0][i + r * I0], w[r], acc);
            const long e = fib * I0 + i;
            if (MODE == 0) out[e] = acc;
            if (MODE == 1) {
                double d = (double)X[e] - (double)acc;
                double y = d * d - c1, t = s1 + y;
                c1 = (t - s1) - y; s1 = t;
                double x2 = (double)X[e] * (double)X[e];
                y = x2 - c2; t = s2 + y;
                c2 = (t - s2) - y; s2 = t;
            }
            if (MODE == 2) {
                out[e] = acc;
                uint64_t n = (uint64_t)e;
                uint64_t blk = (2 * n) >> 2;
                uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), strm, 0u};
                philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
                int wd = (int)((2 * n) & 3);
                double u1 = (double)word_to_uniform(c[wd]), u2 = (double)word_to_uniform(c[wd + 1]);
                double g = (double)(float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
                double y = (double)acc * (double)acc - c1, t = s1 + y;
                c1 = (t - s1) - y; s1 = t;
                y = g * g - c2; t = s2 + y;
                c2 = (t - s2) - y; s2 = t;
            }
        }
    }
    if (MODE != 0) {
        s1 = block_sum_d(s1, sh);
        s2 = block_sum_d(s2, sh);
        if (threadIdx.x == 0) {
            part[2 * blockIdx.x] = s1;
            part[2 * blockIdx.x + 1] = s2;
        }
    }
}

// X += scale * E with scale = eta * sqrt(sum1 / sum2) computed from the partials
__global__ void k_add_noise(float *__restrict__ X, long n, const double *__restrict__ part, int nparts, double eta,
                            uint64_t seed, uint32_t strm) {
    __shared__ double sh[256];
    __shared__ float scale;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    a = block_sum_d(a, sh);
    b = block_sum_d(b, sh);
    if (threadIdx.x == 0) scale = (b > 0.0) ? (float)(eta * sqrt(a / b)) : 0.0f;
    __syncthreads();
    const float sc = scale;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        uint64_t nn = (uint64_t)e;
        uint64_t blk = (2 * nn) >> 2;
        uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), strm, 0u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        int wd = (int)((2 * nn) & 3);
        double u1 = (double)word_to_uniform(c[wd]), u2 = (double)word_to_uniform(c[wd + 1]);
        float g = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
        X[e] = X[e] + sc * g;
    }
}

// ||X - X1||^2 and ||X||^2 (FP64, compensated) partials (src/rse.cu:4-15 without clobbering X1)
__global__ __launch_bounds__(256) void k_rse_partial(const float *__restrict__ X, const float *__restrict__ X1, long n,
                                                     double *__restrict__ part) {
    __shared__ double sh[256];
    double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        double x = X[e];
        double d = x - (double)X1[e];
        double y = d * d - c1, t = s1 + y;
        c1 = (t - s1) - y; s1 = t;
        y = x * x - c2; t = s2 + y;
        c2 = (t - s2) - y; s2 = t;
    }
    s1 = block_sum_d(s1, sh);
    s2 = block_sum_d(s2, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

__global__ void k_ratio(const double *__restrict__ part, int nparts, double *__restrict__ out) {
    __shared__ double sh[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    a = block_sum_d(a, sh);
    b = block_sum_d(b, sh);
    if (threadIdx.x == 0) out[0] = (b > 0.0) ? sqrt(a / b) : 0.0;
}

// Generic mode-n product on a (P, In, Q) view:
//   Y[p + a*P + q*P*Ra] = sum_{i < In} U[a*lda + i*ldi] * X[p + i*P + q*P*In]
// (reference ttm, src/ttm.cu:21,29,37, with lda = 1, ldi = d; transposed factor with
//  lda = In, ldi = 1; Tucker expansion core -> tensor with lda = 1, ldi = I).
__device__ __forceinline__ float ld_elem(const float *p, long i) { return p[i]; }
__device__ __forceinline__ float ld_elem(const uint16_t *p, long i) { return __uint_as_float((uint32_t)p[i] << 16); }
// the factor entry as it enters the product: unchanged for FP32 storage, rounded to BF16
// (nearest-even) for BF16 storage, as on the MFMA path
__device__ __forceinline__ float coef(float u, const float *) { return u; }
__device__ __forceinline__ float coef(float u, const uint16_t *) {
    return __uint_as_float(f32_to_bf16_bits(u) << 16);
}

template <typename T>
__global__ __launch_bounds__(256) void k_ttm_generic(const T *__restrict__ X, long P, long In, long Q,
                                                     const float *__restrict__ U, long lda, long ldi, long Ra,
                                                     float *__restrict__ Y) {
    const long n = P * Ra * Q;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const long p = e % P;
        const long a = (e / P) % Ra;
        const long q = e / (P * Ra);
        const T *x = X + p + q * P * In;
        const float *u = U + a * lda;
        float acc = 0.0f;
        for (long i = 0; i < In; ++i) acc = fmaf(coef(u[i * ldi], x), ld_elem(x, i * P), acc);
        Y[e] = acc;
    }
}

// X is FP32 (xtype TDX_F32) or BF16 bit patterns (TDX_BF16)
int ttm_mfma_launch(const void *X, int xtype, long P, long In, long Q, const float *U, long lda, long ldi, long Ra,
                    float *Y, hipStream_t s);

int ttm_generic_t(const void *X, int xtype, int order, const long *dims, int mode, const float *U, long lda, long ldi,
                  long Ra, float *Y, hipStream_t s) {
    long P = 1, Q = 1;
    for (int m = 0; m < mode; ++m) P *= dims[m];
    for (int m = mode + 1; m < order; ++m) Q *= dims[m];
    // MFMA kernel for the shapes it covers, the per-output kernel otherwise
    const int mf = ttm_mfma_launch(X, xtype, P, dims[mode], Q, U, lda, ldi, Ra, Y, s);
    if (mf < 0) return record(TDX_ERR_HIP);
    if (mf > 0) return TDX_OK;
    const long n = P * Ra * Q;
    long g = ceil_div(n, 256);
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    if (xtype == TDX_BF16)
        k_ttm_generic<uint16_t><<<(int)g, 256, 0, s>>>((const uint16_t *)X, P, dims[mode], Q, U, lda, ldi, Ra, Y);
    else
        k_ttm_generic<float><<<(int)g, 256, 0, s>>>((const float *)X, P, dims[mode], Q, U, lda, ldi, Ra, Y);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

int ttm_generic(const float *X, int order, const long *dims, int mode, const float *U, long lda, long ldi, long Ra,
                float *Y, hipStream_t s) {
    return ttm_generic_t(X, TDX_F32, order, dims, mode, U, lda, ldi, Ra, Y, s);
}

static int cp_fibres_launch(int mode, int order, const long *dims, const float *const *U, long R, float *out,
                            const float *X, double *part, int blocks, uint64_t seed, uint32_t strm, hipStream_t s) {
    if (R > 256) return record(TDX_ERR_UNSUPPORTED);
    FacArgs f{};
    f.order = order;
    f.R = (int)R;
    for (int m = 0; m < 4; ++m) {
        f.U[m] = m < order ? U[m] : nullptr;
        f.dims[m] = m < order ? dims[m] : 1;
    }
    if (mode == 0) k_cp_fibres<0><<<blocks, 256, 0, s>>>(f, out, X, part, seed, strm);
    else if (mode == 1) k_cp_fibres<1><<<blocks, 256, 0, s>>>(f, out, X, part, seed, strm);
    else k_cp_fibres<2><<<blocks, 256, 0, s>>>(f, out, X, part, seed, strm);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

}  // namespace tdx

using namespace tdx;

extern "C" {

int tdx_gencp(float *rec, const float *A, const float *B, const float *C, long a, long b, long c, long r) {
    TDX_CHECK_ARG(rec && A && B && C && a > 0 && b > 0 && c > 0 && r > 0);
    const long dims[3] = {a, b, c};
    const float *U[3] = {A, B, C};
    return record(cp_fibres_launch(0, 3, dims, U, r, rec, nullptr, nullptr, 2048, 0, 0, stream()));
}

int tdx_gencp4(float *rec, const float *A, const float *B, const float *C, const float *D, long a, long b, long c,
               long d, long r) {
    TDX_CHECK_ARG(rec && A && B && C && D && a > 0 && b > 0 && c > 0 && d > 0 && r > 0);
    const long dims[4] = {a, b, c, d};
    const float *U[4] = {A, B, C, D};
    return record(cp_fibres_launch(0, 4, dims, U, r, rec, nullptr, nullptr, 2048, 0, 0, stream()));
}

int tdx_cp_relerr(const void *X, int dtype, const float *scales, int order, const long *dims,
                  const float *const *factors, long rank, double *relerr) {
    TDX_CHECK_ARG(X && dims && factors && relerr && order >= 2 && order <= 4 && rank > 0);
    if (dtype != TDX_F32) return record(TDX_ERR_UNSUPPORTED);
    (void)scales;
    const int blocks = 2048;
    void *ws = nullptr;
    int st = workspace(WS_MISC0, sizeof(double) * (2 * blocks + 1), &ws);
    if (st) return st;
    double *part = (double *)ws;
    hipStream_t s = stream();
    st = cp_fibres_launch(1, order, dims, factors, rank, nullptr, (const float *)X, part, blocks, 0, 0, s);
    if (st) return record(st);
    k_ratio<<<1, 256, 0, s>>>(part, blocks, part + 2 * blocks);
    TDX_LAUNCH_CHECK();
    TDX_HIP(hipMemcpyAsync(relerr, part + 2 * blocks, sizeof(double), hipMemcpyDeviceToHost, s));
    TDX_HIP(hipStreamSynchronize(s));
    return record(TDX_OK);
}

int tdx_gen_cp_tensor(float *X, int order, const long *dims, const float *const *factors, long rank, double eta,
                      uint64_t seed, uint32_t strm) {
    TDX_CHECK_ARG(X && dims && factors && order >= 2 && order <= 4 && rank > 0 && eta >= 0.0);
    const int blocks = 2048;
    void *ws = nullptr;
    int st = workspace(WS_MISC1, sizeof(double) * 2 * blocks, &ws);
    if (st) return st;
    double *part = (double *)ws;
    hipStream_t s = stream();
    long N = 1;
    for (int m = 0; m < order; ++m) N *= dims[m];
    st = cp_fibres_launch(2, order, dims, factors, rank, X, nullptr, part, blocks, seed, strm, s);
    if (st) return record(st);
    long g = ceil_div(N, 256);
    if (g > 65536) g = 65536;
    k_add_noise<<<(int)g, 256, 0, s>>>(X, N, part, blocks, eta, seed, strm);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_rse(const float *X, const float *X1, long m, float *error) {
    TDX_CHECK_ARG(X && X1 && error && m > 0);
    const int blocks = 1024;
    void *ws = nullptr;
    int st = workspace(WS_MISC0, sizeof(double) * (2 * blocks + 1), &ws);
    if (st) return st;
    double *part = (double *)ws;
    hipStream_t s = stream();
    k_rse_partial<<<blocks, 256, 0, s>>>(X, X1, m, part);
    k_ratio<<<1, 256, 0, s>>>(part, blocks, part + 2 * blocks);
    TDX_LAUNCH_CHECK();
    double r = 0.0;
    TDX_HIP(hipMemcpyAsync(&r, part + 2 * blocks, sizeof(double), hipMemcpyDeviceToHost, s));
    TDX_HIP(hipStreamSynchronize(s));
    *error = (float)r;
    return record(TDX_OK);
}

int tdx_ttm(const float *X, const float *U, float *XU, long flag, long a, long b, long c, long d) {
    TDX_CHECK_ARG(X && U && XU && a > 0 && b > 0 && c > 0 && d > 0);
    TDX_CHECK_ARG(flag >= 1 && flag <= 3);
    const long dims[3] = {a, b, c};
    // U is d x (mode size), col-major: U[p + i*d]
    return record(ttm_generic(X, 3, dims, (int)flag - 1, U, 1, d, d, XU, stream()));
}

int tdx_ttm_t(const float *X, int order, const long *dims, int mode, const float *U, long r, float *Y) {
    TDX_CHECK_ARG(X && dims && U && Y && order >= 2 && order <= 4 && mode >= 0 && mode < order && r > 0);
    return record(ttm_generic(X, order, dims, mode, U, dims[mode], 1, r, Y, stream()));
}

int tdx_ttm_ex(const void *X, int dtype, int order, const long *dims, int mode, const float *U, long r, float *Y) {
    TDX_CHECK_ARG(X && dims && U && Y && order >= 2 && order <= 4 && mode >= 0 && mode < order && r > 0);
    TDX_CHECK_ARG(dtype == TDX_F32 || dtype == TDX_BF16);
    for (int m = 0; m < order; ++m) TDX_CHECK_ARG(dims[m] > 0);
    return record(ttm_generic_t(X, dtype, order, dims, mode, U, dims[mode], 1, r, Y, stream()));
}

// Tucker reconstruction core x_1 U_1 ... x_d U_d by successive expansions
static int gentucker_n(float *rec, const float *core, const float *const *U, int order, const long *dims,
                       const long *ranks) {
    hipStream_t s = stream();
    long cur[4];
    for (int m = 0; m < order; ++m) cur[m] = ranks[m];
    // the intermediates of steps 0 .. order-2 ping-pong between two workspaces; the last
    // step writes into rec, so it does not size them
    long maxsz = 1;
    {
        long sz = 1;
        for (int m = 0; m < order; ++m) sz *= ranks[m];
        for (int m = 0; m < order - 1; ++m) {
            sz = sz / ranks[m] * dims[m];
            if (sz > maxsz) maxsz = sz;
        }
    }
    void *w0 = nullptr, *w1 = nullptr;
    int st = workspace(WS_MISC2, sizeof(float) * maxsz, &w0);
    if (st) return st;
    st = workspace(WS_MISC3, sizeof(float) * maxsz, &w1);
    if (st) return st;
    const float *src = core;
    for (int m = 0; m < order; ++m) {
        float *dst = (m == order - 1) ? rec : ((m & 1) ? (float *)w1 : (float *)w0);
        // Y[p, i, q] = sum_c U_m[i + c*I_m] * src[p, c, q]
        st = ttm_generic(src, order, cur, m, U[m], 1, dims[m], dims[m], dst, s);
        if (st) return st;
        cur[m] = dims[m];
        src = dst;
    }
    return TDX_OK;
}

int tdx_gentucker(float *rec, const float *core, const float *A, const float *B, const float *C, long a, long b,
                  long c, long r1, long r2, long r3) {
    TDX_CHECK_ARG(rec && core && A && B && C && a > 0 && b > 0 && c > 0 && r1 > 0 && r2 > 0 && r3 > 0);
    const long dims[3] = {a, b, c}, ranks[3] = {r1, r2, r3};
    const float *U[3] = {A, B, C};
    return record(gentucker_n(rec, core, U, 3, dims, ranks));
}

int tdx_gentucker4(float *rec, const float *core, const float *A, const float *B, const float *C, const float *D,
                   long a, long b, long c, long d, long r1, long r2, long r3, long r4) {
    TDX_CHECK_ARG(rec && core && A && B && C && D && a > 0 && b > 0 && c > 0 && d > 0);
    TDX_CHECK_ARG(r1 > 0 && r2 > 0 && r3 > 0 && r4 > 0);
    const long dims[4] = {a, b, c, d}, ranks[4] = {r1, r2, r3, r4};
    const float *U[4] = {A, B, C, D};
    return record(gentucker_n(rec, core, U, 4, dims, ranks));
}

}  // extern "C"
