// mlp.hpp -- the small MLP of the learned ALS extrapolation step.
//
// y = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3, GELU in its exact erf form
// (Hendrycks & Gimpel 2016).  Parameters live in one FP32 device buffer:
//   [W1 (width x nin, row-major) | b1 (width) | W2 (width x width) | b2 | W3 (nout x width) | b3]
#pragma once
#include "common.hpp"

namespace tdx {

struct MlpShape {
    long nin, width, nout;
    __host__ __device__ long n_w1() const { return width * nin; }
    __host__ __device__ long n_w2() const { return width * width; }
    __host__ __device__ long n_w3() const { return nout * width; }
    __host__ __device__ long off_b1() const { return n_w1(); }
    __host__ __device__ long off_w2() const { return off_b1() + width; }
    __host__ __device__ long off_b2() const { return off_w2() + n_w2(); }
    __host__ __device__ long off_w3() const { return off_b2() + width; }
    __host__ __device__ long off_b3() const { return off_w3() + n_w3(); }
    __host__ __device__ long count() const { return off_b3() + nout; }
};

__device__ __forceinline__ float gelu_f(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_df(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * __expf(-0.5f * x * x);
}

// Forward pass for ONE sample inside a workgroup of >= width threads (width <= 128).
// x: nin inputs (any memory space); returns y[0] in all threads.  sh needs 2*width floats.
__device__ inline float mlp_forward_one(const float *__restrict__ P, MlpShape s, const float *x, float *sh) {
    const int t = threadIdx.x;
    float *h1 = sh, *h2 = sh + s.width;
    if (t < s.width) {
        float z = P[s.off_b1() + t];
        for (int i = 0; i < s.nin; ++i) z = fmaf(P[t * s.nin + i], x[i], z);
        h1[t] = gelu_f(z);
    }
    __syncthreads();
    if (t < s.width) {
        float z = P[s.off_b2() + t];
        for (int i = 0; i < s.width; ++i) z = fmaf(P[s.off_w2() + t * s.width + i], h1[i], z);
        h2[t] = gelu_f(z);
    }
    __syncthreads();
    float y = P[s.off_b3()];
    for (int i = 0; i < s.width; ++i) y = fmaf(P[s.off_w3() + i], h2[i], y);
    __syncthreads();
    return y;
}

}  // namespace tdx
