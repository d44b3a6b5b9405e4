// elementwise.hip -- permutation, elementwise and format-conversion kernels.
// All are HBM-bound streaming kernels: grid-stride loops, 64-bit indices, vector
// loads where the extent allows it, LDS-tiled transposes for the unfoldings.
#include "common.hpp"

namespace tdx {

static constexpr int kBlock = 256;

static inline int grid_for(long n, int per_thread = 1) {
    long g = ceil_div(n, (long)kBlock * per_thread);
    if (g < 1) g = 1;
    if (g > 65536L * 8) g = 65536L * 8;
    return (int)g;
}

// ---------------------------------------------------------------------------
// elementwise products (src/hadmard.cu:2-30, src/kernel.cu:3-9)
__global__ void k_hadamard(const float *__restrict__ A, const float *__restrict__ B,
                           float *__restrict__ C, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        C[i] = A[i] * B[i];
}

__global__ void k_hadamard3(const float *__restrict__ A, const float *__restrict__ B,
                            const float *__restrict__ C, float *__restrict__ D, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float t = A[i] * B[i];
        D[i] = t * C[i];
    }
}

// right := left - right (src/tminust.cu:2-8)
__global__ void k_tminust(const float *__restrict__ L, float *__restrict__ R, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        R[i] = L[i] - R[i];
}

__global__ void k_identity(float *__restrict__ A, long m) {
    long n = m * m;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        long r = e % m, c = e / m;
        A[e] = (r == c) ? 1.0f : 0.0f;
    }
}

// Khatri-Rao product, column t of the result is kron(A(:,t), B(:,t)) with the B index
// fastest (src/kr.cu:9): out[jb + ia*b + t*a*b] = A[ia + t*a] * B[jb + t*b]
__global__ void k_kr(const float *__restrict__ A, const float *__restrict__ B, float *__restrict__ O,
                     long a, long b, long r) {
    long n = a * b * r;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        long jb = e % b;
        long rest = e / b;
        long ia = rest % a;
        long t = rest / a;
        O[e] = A[ia + t * a] * B[jb + t * b];
    }
}

// Kronecker product (src/tensorproduct.cu:9-11): A (a x b) (x) B (c x d) -> (a c) x (b d)
__global__ void k_kron(const float *__restrict__ A, const float *__restrict__ B, float *__restrict__ C,
                       long a, long b, long c, long d) {
    long rows = a * c, n = rows * b * d;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        long row = e % rows, col = e / rows;
        long ia = row / c, ic = row % c;
        long jb = col / d, jd = col % d;
        C[e] = A[ia + jb * a] * B[ic + jd * c];
    }
}

// ---------------------------------------------------------------------------
// Batched transpose through a padded 64x64 LDS tile: in is rows x cols (col-major,
// element (r,c) at r + c*rows) per batch, out is cols x rows.  Serves the mode-2
// unfolding (batch = k slices of m x n) and the mode-3 unfolding (one (m n) x k matrix).
template <typename T>
__global__ __launch_bounds__(256) void k_transpose(const T *__restrict__ in, T *__restrict__ out,
                                                   long rows, long cols, long batch) {
    __shared__ T tile[64][65];
    long tr = ceil_div(rows, 64), tc = ceil_div(cols, 64);
    long ntiles = tr * tc * batch;
    int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        long bt = t / (tr * tc), rem = t % (tr * tc);
        long r0 = (rem % tr) * 64, c0 = (rem / tr) * 64;
        const T *src = in + bt * rows * cols;
        T *dst = out + bt * rows * cols;
        for (int cc = ty; cc < 64; cc += 4) {
            long r = r0 + tx, c = c0 + cc;
            if (r < rows && c < cols) tile[cc][tx] = src[r + c * rows];
        }
        __syncthreads();
        for (int rr = ty; rr < 64; rr += 4) {
            long c = c0 + tx, r = r0 + rr;
            if (r < rows && c < cols) dst[c + r * cols] = tile[tx][rr];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void k_copy(const T *__restrict__ in, T *__restrict__ out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// ---------------------------------------------------------------------------
// formats
__global__ void k_f2h(const float *__restrict__ A, uint16_t *__restrict__ B, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        _Float16 h = (_Float16)A[i];  // v_cvt_f16_f32, round to nearest even
        B[i] = __builtin_bit_cast(uint16_t, h);
    }
}

__global__ void k_to_bf16(const float *__restrict__ X, uint16_t *__restrict__ Q, long n) {
    long n4 = n >> 2;
    const float4 *X4 = reinterpret_cast<const float4 *>(X);
    uint2 *Q4 = reinterpret_cast<uint2 *>(Q);
    bool vec = ((((uintptr_t)X) | ((uintptr_t)Q)) & 15) == 0;
    long start = blockIdx.x * (long)blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    if (vec) {
        for (long i = start; i < n4; i += stride) {
            float4 v = X4[i];
            uint2 o;
            o.x = (uint32_t)f32_to_bf16_bits(v.x) | ((uint32_t)f32_to_bf16_bits(v.y) << 16);
            o.y = (uint32_t)f32_to_bf16_bits(v.z) | ((uint32_t)f32_to_bf16_bits(v.w) << 16);
            Q4[i] = o;
        }
        for (long i = 4 * n4 + start; i < n; i += stride) Q[i] = f32_to_bf16_bits(X[i]);
    } else {
        for (long i = start; i < n; i += stride) Q[i] = f32_to_bf16_bits(X[i]);
    }
}

__global__ void k_from_bf16(const uint16_t *__restrict__ Q, float *__restrict__ X, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        X[i] = __uint_as_float((uint32_t)Q[i] << 16);
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, long n, uint32_t v) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// per-mode-1-slice absolute maximum: |x| compared as unsigned bit patterns (order independent)
__global__ void k_slice_amax(const float *__restrict__ X, uint32_t *__restrict__ amax, long I, long rest) {
    // block: 256 threads = 64 slices x 4 column groups; grid-stride over column chunks
    long i = blockIdx.x * 64L + (threadIdx.x & 63);
    int cg = threadIdx.x >> 6;
    if (i >= I) return;
    uint32_t m = 0;
    for (long t = blockIdx.y * 4L + cg; t < rest; t += 4L * gridDim.y) {
        uint32_t u = __float_as_uint(X[i + t * I]) & 0x7fffffffu;
        if (u > 0x7f800000u) u = 0u;  // NaN does not enter the scale (same as the oracle)
        m = u > m ? u : m;
    }
    atomicMax(&amax[i], m);
}

__global__ void k_scales(const uint32_t *__restrict__ amax, float *__restrict__ scales, long I) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < I; i += (long)gridDim.x * blockDim.x) {
        float a = __uint_as_float(amax[i]);
        scales[i] = (a > 0.0f) ? a / 448.0f : 1.0f;
    }
}

__global__ void k_to_e4m3(const float *__restrict__ X, const float *__restrict__ scales, uint8_t *__restrict__ Q,
                          long I, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        Q[e] = f32_to_e4m3_bits(X[e] / scales[e % I]);
}

// the same, 16 consecutive elements per thread (I % 16 == 0, so they are 16 consecutive
// slices): four 16-byte nontemporal loads, one 16-byte store; X, Q 16-byte aligned
__global__ __launch_bounds__(256) void k_to_e4m3_16(const float *__restrict__ X, const float *__restrict__ scales,
                                                    uint8_t *__restrict__ Q, long I, long n16) {
    const float4 *X4 = reinterpret_cast<const float4 *>(X);
    for (long b = blockIdx.x * 256L + threadIdx.x; b < n16; b += (long)gridDim.x * 256) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ld_stream(X4 + 4 * b + u);
        const float4 *sc = reinterpret_cast<const float4 *>(scales + (16 * b) % I);
        uint32_t w[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float4 c = sc[u];
            w[u] = (uint32_t)f32_to_e4m3_bits(v[u].x / c.x) | ((uint32_t)f32_to_e4m3_bits(v[u].y / c.y) << 8) |
                   ((uint32_t)f32_to_e4m3_bits(v[u].z / c.z) << 16) | ((uint32_t)f32_to_e4m3_bits(v[u].w / c.w) << 24);
        }
        typedef unsigned v4u __attribute__((ext_vector_type(4)));
        __builtin_nontemporal_store(v4u{w[0], w[1], w[2], w[3]}, reinterpret_cast<v4u *>(Q) + b);
    }
}

// slice maxima, 4 consecutive slices per lane (I % 4 == 0, X 16-byte aligned): a wave reads
// 1 KB of one i_rest step per load, 4 steps in flight
__global__ __launch_bounds__(256) void k_slice_amax4(const float *__restrict__ X, uint32_t *__restrict__ amax, long I,
                                                     long rest) {
    const long I4 = I / 4;
    const long q = blockIdx.x * 64L + (threadIdx.x & 63);
    const int cg = threadIdx.x >> 6;
    if (q >= I4) return;
    const float4 *X4 = reinterpret_cast<const float4 *>(X);
    uint32_t m[4] = {0u, 0u, 0u, 0u};
    auto take = [&](const float4 &v) {
        const float w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t u = __float_as_uint(w[k]) & 0x7fffffffu;
            if (u > 0x7f800000u) u = 0u;
            m[k] = u > m[k] ? u : m[k];
        }
    };
    const long step = 4L * gridDim.y;
    long t = blockIdx.y * 4L + cg;
    for (; t + 3 * step < rest; t += 4 * step) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ld_stream(X4 + q + (t + u * step) * I4);
#pragma unroll
        for (int u = 0; u < 4; ++u) take(v[u]);
    }
    for (; t < rest; t += step) take(ld_stream(X4 + q + t * I4));
#pragma unroll
    for (int k = 0; k < 4; ++k) atomicMax(&amax[4 * q + k], m[k]);
}

__device__ __forceinline__ float e4m3_to_f32(uint8_t q) {
    uint32_t e = (q >> 3) & 0xf, m = q & 7;
    float s = (q & 0x80) ? -1.0f : 1.0f;
    if (e == 0xf && m == 7) return __uint_as_float(0x7fc00000u);
    if (e == 0) return s * (float)m * (1.0f / 512.0f);
    return s * __uint_as_float(((e + 120u) << 23) | (m << 20));
}

__global__ void k_from_e4m3(const uint8_t *__restrict__ Q, const float *__restrict__ scales, float *__restrict__ X,
                            long I, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        X[e] = scales[e % I] * e4m3_to_f32(Q[e]);
}

// ---------------------------------------------------------------------------
// Philox streams (identical definition to the oracle)
__global__ void k_fill_uniform(float *__restrict__ out, long n, uint64_t seed, uint32_t strm) {
    long nb = (n + 3) >> 2;
    for (long b = blockIdx.x * (long)blockDim.x + threadIdx.x; b < nb; b += (long)gridDim.x * blockDim.x) {
        uint32_t c[4] = {(uint32_t)b, (uint32_t)((uint64_t)b >> 32), strm, 0u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            long i = 4 * b + w;
            if (i < n) out[i] = word_to_uniform(c[w]);
        }
    }
}

__device__ __forceinline__ double normal_at(uint64_t seed, uint32_t strm, uint64_t n) {
    // uniforms 2n and 2n+1 share one Philox block when n is aligned (2n even)
    uint64_t blk = (2 * n) >> 2;
    uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), strm, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    int w = (int)((2 * n) & 3);
    double u1 = (double)word_to_uniform(c[w]);
    double u2 = (double)word_to_uniform(c[w + 1]);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ void k_fill_normal(float *__restrict__ out, long n, uint64_t seed, uint32_t strm) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)normal_at(seed, strm, (uint64_t)i);
}

}  // namespace tdx

using namespace tdx;

extern "C" {

int tdx_hadamard(const float *A, const float *B, float *C, long a, long b) {
    TDX_CHECK_ARG(A && B && C && a >= 0 && b >= 0);
    long n = a * b;
    if (n == 0) return record(TDX_OK);
    k_hadamard<<<grid_for(n), kBlock, 0, stream()>>>(A, B, C, n);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_hadamard3(const float *A, const float *B, const float *C, float *ABC, long a, long b) {
    TDX_CHECK_ARG(A && B && C && ABC && a >= 0 && b >= 0);
    long n = a * b;
    if (n == 0) return record(TDX_OK);
    k_hadamard3<<<grid_for(n), kBlock, 0, stream()>>>(A, B, C, ABC, n);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_kr(const float *A, const float *B, float *AkrB, long a, long b, long r) {
    TDX_CHECK_ARG(A && B && AkrB && a >= 0 && b >= 0 && r >= 0);
    long n = a * b * r;
    if (n == 0) return record(TDX_OK);
    k_kr<<<grid_for(n), kBlock, 0, stream()>>>(A, B, AkrB, a, b, r);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_tensor_product(const float *A, const float *B, float *C, long a, long b, long c, long d) {
    TDX_CHECK_ARG(A && B && C && a >= 0 && b >= 0 && c >= 0 && d >= 0);
    long n = a * b * c * d;
    if (n == 0) return record(TDX_OK);
    k_kron<<<grid_for(n), kBlock, 0, stream()>>>(A, B, C, a, b, c, d);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_t2m(const float *T1, float *T2, long flag, long m, long n, long k) {
    TDX_CHECK_ARG(T1 && T2 && m > 0 && n > 0 && k > 0);
    TDX_CHECK_ARG(flag >= 1 && flag <= 3);
    long N = m * n * k;
    if (flag == 1) {
        k_copy<float><<<grid_for(N), kBlock, 0, stream()>>>(T1, T2, N);
    } else if (flag == 2) {
        long tiles = ceil_div(m, 64) * ceil_div(n, 64) * k;
        int g = (int)(tiles < 65536L * 4 ? tiles : 65536L * 4);
        k_transpose<float><<<g, 256, 0, stream()>>>(T1, T2, m, n, k);
    } else {
        long tiles = ceil_div(m * n, 64) * ceil_div(k, 64);
        int g = (int)(tiles < 65536L * 4 ? tiles : 65536L * 4);
        k_transpose<float><<<g, 256, 0, stream()>>>(T1, T2, m * n, k, 1);
    }
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_tminust(const float *left, float *right, long m) {
    TDX_CHECK_ARG(left && right && m >= 0);
    if (m == 0) return record(TDX_OK);
    k_tminust<<<grid_for(m), kBlock, 0, stream()>>>(left, right, m);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_init_identity(float *A, long m) {
    TDX_CHECK_ARG(A && m >= 0);
    if (m == 0) return record(TDX_OK);
    k_identity<<<grid_for(m * m), kBlock, 0, stream()>>>(A, m);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_f2h(const float *A, uint16_t *B, long num) {
    TDX_CHECK_ARG(A && B && num >= 0);
    if (num == 0) return record(TDX_OK);
    k_f2h<<<grid_for(num), kBlock, 0, stream()>>>(A, B, num);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_quantize_bf16(const float *X, uint16_t *Xq, long n) {
    TDX_CHECK_ARG(X && Xq && n >= 0);
    if (n == 0) return record(TDX_OK);
    k_to_bf16<<<grid_for(n, 4), kBlock, 0, stream()>>>(X, Xq, n);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_dequantize_bf16(const uint16_t *Xq, float *X, long n) {
    TDX_CHECK_ARG(X && Xq && n >= 0);
    if (n == 0) return record(TDX_OK);
    k_from_bf16<<<grid_for(n), kBlock, 0, stream()>>>(Xq, X, n);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

}  // extern "C"

namespace tdx {

// scales from the slice maxima, then the E4M3 codes
int fp8_from_amax(const float *X, const uint32_t *amax, uint8_t *Xq, float *scales, long I, long n_rest,
                  hipStream_t s) {
    const long n = I * n_rest;
    k_scales<<<grid_for(I), kBlock, 0, s>>>(amax, scales, I);
    const bool vec = I % 16 == 0 && ((((uintptr_t)X) | ((uintptr_t)Xq) | ((uintptr_t)scales)) & 15) == 0;
    if (vec) {
        const long n16 = n / 16;
        const int blocks = (int)std::min<long>(std::max<long>(1, ceil_div(n16, 256L)), 8L * device_cus());
        k_to_e4m3_16<<<blocks, 256, 0, s>>>(X, scales, Xq, I, n16);
    } else {
        k_to_e4m3<<<grid_for(n), kBlock, 0, s>>>(X, scales, Xq, I, n);
    }
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

int slice_amax(const float *X, uint32_t *amax, long I, long n_rest, hipStream_t s) {
    k_fill_u32<<<grid_for(I), kBlock, 0, s>>>(amax, I, 0u);
    long gy = n_rest < 1024 ? n_rest : 1024;
    if (I % 4 == 0 && (((uintptr_t)X) & 15) == 0) {
        dim3 g((unsigned)ceil_div(I / 4, 64L), (unsigned)((gy + 3) / 4));
        k_slice_amax4<<<g, 256, 0, s>>>(X, amax, I, n_rest);
    } else {
        dim3 g((unsigned)ceil_div(I, 64), (unsigned)((gy + 3) / 4));
        k_slice_amax<<<g, 256, 0, s>>>(X, amax, I, n_rest);
    }
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

}  // namespace tdx

extern "C" {

int tdx_quantize_fp8_mode1(const float *X, uint8_t *Xq, float *scales, long I, long n_rest) {
    TDX_CHECK_ARG(X && Xq && scales && I > 0 && n_rest > 0);
    void *ws = nullptr;
    int st = workspace(WS_MISC3, sizeof(uint32_t) * I, &ws);
    if (st) return st;
    uint32_t *amax = (uint32_t *)ws;
    st = slice_amax(X, amax, I, n_rest, stream());
    if (!st) st = fp8_from_amax(X, amax, Xq, scales, I, n_rest, stream());
    return record(st);
}

int tdx_dequantize_fp8_mode1(const uint8_t *Xq, const float *scales, float *X, long I, long n_rest) {
    TDX_CHECK_ARG(X && Xq && scales && I > 0 && n_rest > 0);
    k_from_e4m3<<<grid_for(I * n_rest), kBlock, 0, stream()>>>(Xq, scales, X, I, I * n_rest);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_fill_uniform(float *out, long n, uint64_t seed, uint32_t strm) {
    TDX_CHECK_ARG(out && n >= 0);
    if (n == 0) return record(TDX_OK);
    k_fill_uniform<<<grid_for(ceil_div(n, 4)), kBlock, 0, stream()>>>(out, n, seed, strm);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_fill_normal(float *out, long n, uint64_t seed, uint32_t strm) {
    TDX_CHECK_ARG(out && n >= 0);
    if (n == 0) return record(TDX_OK);
    k_fill_normal<<<grid_for(n), kBlock, 0, stream()>>>(out, n, seed, strm);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

}  // extern "C"
