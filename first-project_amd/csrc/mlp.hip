// mlp.hip -- forward, backward and Adam kernels of the extrapolation-step MLP (FP32).
// All reductions are loops in a fixed order (bit-reproducible gradients).
#include "mlp.hpp"
#include "linalg.hpp"
#include <math.h>
#include <random>
#include <vector>

namespace tdx {

// out[s, o] = act(b[o] + sum_i W[o, i] in[s, i]); z (pre-activation) kept when non-null
__global__ void k_dense(const float *__restrict__ in, const float *__restrict__ W, const float *__restrict__ b,
                        float *__restrict__ z, float *__restrict__ out, long batch, long nin, long nout, int act) {
    const long n = batch * nout;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long s = e / nout, o = e % nout;
        float acc = b[o];
        const float *w = W + o * nin;
        const float *x = in + s * nin;
        for (long i = 0; i < nin; ++i) acc = fmaf(w[i], x[i], acc);
        if (z) z[e] = acc;
        out[e] = act ? gelu_f(acc) : acc;
    }
}

// dy = (y - t) / (batch * nout); sq = (y - t)^2
__global__ void k_out_grad(const float *__restrict__ y, const float *__restrict__ t, float *__restrict__ dy,
                           float *__restrict__ sq, long n, float inv) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const float d = y[e] - t[e];
        dy[e] = d * inv;
        sq[e] = d * d;
    }
}

__global__ void k_loss(const float *__restrict__ sq, long n, float inv2, float *__restrict__ loss) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long e = threadIdx.x; e < n; e += 256) s += (double)sq[e];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *loss = (float)(s * (double)inv2);
}

// gW[o, i] = sum_s d[s, o] in[s, i]; gb[o] = sum_s d[s, o]
__global__ void k_wgrad(const float *__restrict__ d, const float *__restrict__ in, float *__restrict__ gW,
                        float *__restrict__ gb, long batch, long nin, long nout) {
    const long n = nout * nin + nout;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        float acc = 0.0f;
        if (e < nout * nin) {
            const long o = e / nin, i = e % nin;
            for (long s = 0; s < batch; ++s) acc = fmaf(d[s * nout + o], in[s * nin + i], acc);
            gW[e] = acc;
        } else {
            const long o = e - nout * nin;
            for (long s = 0; s < batch; ++s) acc += d[s * nout + o];
            gb[o] = acc;
        }
    }
}

// dprev[s, i] = (sum_o W[o, i] d[s, o]) * gelu'(zprev[s, i])
__global__ void k_dgrad(const float *__restrict__ d, const float *__restrict__ W, const float *__restrict__ zprev,
                        float *__restrict__ dprev, long batch, long nin, long nout) {
    const long n = batch * nin;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long s = e / nin, i = e % nin;
        float acc = 0.0f;
        for (long o = 0; o < nout; ++o) acc = fmaf(W[o * nin + i], d[s * nout + o], acc);
        dprev[e] = acc * gelu_df(zprev[e]);
    }
}

// Adam (Kingma & Ba 2015, Algorithm 1); c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const float gg = g[e];
        const float mm = b1 * m[e] + (1.0f - b1) * gg;
        const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
        m[e] = mm;
        v[e] = vv;
        p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    }
}

__global__ void k_glorot(float *__restrict__ p, long n, float a, uint64_t seed, uint32_t strm) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        p[e] = a * (2.0f * word_to_uniform(philox_word(seed, strm, (uint64_t)e)) - 1.0f);
}

__global__ void k_zero(float *__restrict__ p, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = 0.0f;
}

__global__ void k_gather_rows(const float *__restrict__ src, const int *__restrict__ idx, float *__restrict__ dst,
                              long rows, long cols) {
    const long n = rows * cols;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        dst[e] = src[(long)idx[e / cols] * cols + e % cols];
}

static int grid1(long n) {
    long g = ceil_div(n, 256);
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (int)g;
}

struct MlpWs {
    float *z1 = nullptr, *h1 = nullptr, *z2 = nullptr, *h2 = nullptr, *y = nullptr, *dy = nullptr, *d2 = nullptr,
          *d1 = nullptr, *sq = nullptr;
    ~MlpWs() {
        float *ps[] = {z1, h1, z2, h2, y, dy, d2, d1, sq};
        for (float *p : ps)
            if (p) (void)hipFree(p);
    }
    int alloc(long batch, const MlpShape &s) {
        long bw = batch * s.width, bo = batch * s.nout;
        float **ps[] = {&z1, &h1, &z2, &h2, &d2, &d1};
        for (float **p : ps) TDX_HIP(hipMalloc(p, sizeof(float) * bw));
        TDX_HIP(hipMalloc(&y, sizeof(float) * bo));
        TDX_HIP(hipMalloc(&dy, sizeof(float) * bo));
        TDX_HIP(hipMalloc(&sq, sizeof(float) * bo));
        return TDX_OK;
    }
};

static int forward(const float *P, const MlpShape &s, const float *x, long batch, MlpWs &w, float *y, hipStream_t st) {
    k_dense<<<grid1(batch * s.width), 256, 0, st>>>(x, P, P + s.off_b1(), w.z1, w.h1, batch, s.nin, s.width, 1);
    k_dense<<<grid1(batch * s.width), 256, 0, st>>>(w.h1, P + s.off_w2(), P + s.off_b2(), w.z2, w.h2, batch, s.width,
                                                   s.width, 1);
    k_dense<<<grid1(batch * s.nout), 256, 0, st>>>(w.h2, P + s.off_w3(), P + s.off_b3(), nullptr, y, batch, s.width,
                                                  s.nout, 0);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

static int backward(const float *P, const MlpShape &s, const float *x, const float *t, long batch, MlpWs &w,
                    float *G, float *loss, hipStream_t st) {
    int rc = forward(P, s, x, batch, w, w.y, st);
    if (rc) return rc;
    const long bo = batch * s.nout;
    const float inv = 1.0f / (float)bo;
    k_out_grad<<<grid1(bo), 256, 0, st>>>(w.y, t, w.dy, w.sq, bo, inv);
    if (loss) k_loss<<<1, 256, 0, st>>>(w.sq, bo, 0.5f * inv, loss);
    k_wgrad<<<grid1(s.n_w3() + s.nout), 256, 0, st>>>(w.dy, w.h2, G + s.off_w3(), G + s.off_b3(), batch, s.width,
                                                     s.nout);
    k_dgrad<<<grid1(batch * s.width), 256, 0, st>>>(w.dy, P + s.off_w3(), w.z2, w.d2, batch, s.width, s.nout);
    k_wgrad<<<grid1(s.n_w2() + s.width), 256, 0, st>>>(w.d2, w.h1, G + s.off_w2(), G + s.off_b2(), batch, s.width,
                                                      s.width);
    k_dgrad<<<grid1(batch * s.width), 256, 0, st>>>(w.d2, P + s.off_w2(), w.z1, w.d1, batch, s.width, s.width);
    k_wgrad<<<grid1(s.n_w1() + s.width), 256, 0, st>>>(w.d1, x, G, G + s.off_b1(), batch, s.nin, s.width);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

static bool shape_ok(long nin, long width, long nout) {
    return nin > 0 && nin <= 64 && width > 0 && width <= 128 && nout > 0 && nout <= 16;
}

}  // namespace tdx

using namespace tdx;

extern "C" {

long tdx_mlp_param_count(long nin, long width, long nout) {
    if (!shape_ok(nin, width, nout)) return -1;
    return MlpShape{nin, width, nout}.count();
}

int tdx_mlp_init(float *params, long nin, long width, long nout, uint64_t seed) {
    TDX_CHECK_ARG(params && shape_ok(nin, width, nout));
    MlpShape s{nin, width, nout};
    hipStream_t st = stream();
    k_glorot<<<grid1(s.n_w1()), 256, 0, st>>>(params, s.n_w1(), sqrtf(6.0f / (float)(nin + width)), seed, 1);
    k_zero<<<1, 256, 0, st>>>(params + s.off_b1(), width);
    k_glorot<<<grid1(s.n_w2()), 256, 0, st>>>(params + s.off_w2(), s.n_w2(), sqrtf(6.0f / (float)(2 * width)), seed, 2);
    k_zero<<<1, 256, 0, st>>>(params + s.off_b2(), width);
    k_glorot<<<grid1(s.n_w3()), 256, 0, st>>>(params + s.off_w3(), s.n_w3(), sqrtf(6.0f / (float)(width + nout)), seed,
                                              3);
    k_zero<<<1, 256, 0, st>>>(params + s.off_b3(), nout);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

int tdx_mlp_forward(const float *params, const float *x, long batch, long nin, long width, long nout, float *y) {
    TDX_CHECK_ARG(params && x && y && batch > 0 && shape_ok(nin, width, nout));
    MlpShape s{nin, width, nout};
    MlpWs w;
    int rc = w.alloc(batch, s);
    if (!rc) rc = forward(params, s, x, batch, w, y, stream());
    if (!rc) TDX_HIP(hipStreamSynchronize(stream()));
    return record(rc);
}

int tdx_mlp_backward(const float *params, const float *x, const float *t, long batch, long nin, long width, long nout,
                     float *grad, float *loss) {
    TDX_CHECK_ARG(params && x && t && grad && batch > 0 && shape_ok(nin, width, nout));
    MlpShape s{nin, width, nout};
    MlpWs w;
    int rc = w.alloc(batch, s);
    if (!rc) rc = backward(params, s, x, t, batch, w, grad, loss, stream());
    if (!rc) TDX_HIP(hipStreamSynchronize(stream()));
    return record(rc);
}

int tdx_mlp_adam(float *params, const float *grad, float *m, float *v, long n, long step, float lr, float b1,
                 float b2, float eps) {
    TDX_CHECK_ARG(params && grad && m && v && n > 0 && step >= 1);
    const float c1 = (float)(1.0 - pow((double)b1, (double)step));
    const float c2 = (float)(1.0 - pow((double)b2, (double)step));
    k_adam<<<grid1(n), 256, 0, stream()>>>(params, grad, m, v, n, lr, b1, b2, eps, c1, c2);
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

// Minibatch Adam training on the device.  X (nsamples x nin) and T (nsamples x nout) are
// device arrays; samples are shuffled every epoch with a seeded host permutation.
// loss_hist (host, may be NULL) receives the last minibatch loss of every epoch.
int tdx_mlp_train(float *params, const float *X, const float *T, long nsamples, long nin, long width, long nout,
                  int epochs, long batch, float lr, uint64_t seed, float *loss_hist) {
    TDX_CHECK_ARG(params && X && T && nsamples > 0 && epochs >= 0 && batch > 0 && shape_ok(nin, width, nout));
    if (batch > nsamples) batch = nsamples;
    MlpShape s{nin, width, nout};
    hipStream_t st = stream();
    MlpWs w;
    int rc = w.alloc(batch, s);
    if (rc) return rc;
    float *G = nullptr, *m = nullptr, *v = nullptr, *xb = nullptr, *tb = nullptr, *loss = nullptr;
    int *idx = nullptr;
    struct Frees {
        void **p[7];
        ~Frees() {
            for (auto q : p)
                if (*q) (void)hipFree(*q);
        }
    } fr{{(void **)&G, (void **)&m, (void **)&v, (void **)&xb, (void **)&tb, (void **)&loss, (void **)&idx}};
    const long n = s.count();
    TDX_HIP(hipMalloc(&G, sizeof(float) * n));
    TDX_HIP(hipMalloc(&m, sizeof(float) * n));
    TDX_HIP(hipMalloc(&v, sizeof(float) * n));
    TDX_HIP(hipMalloc(&xb, sizeof(float) * batch * nin));
    TDX_HIP(hipMalloc(&tb, sizeof(float) * batch * nout));
    TDX_HIP(hipMalloc(&loss, sizeof(float)));
    TDX_HIP(hipMalloc(&idx, sizeof(int) * nsamples));
    k_zero<<<grid1(n), 256, 0, st>>>(m, n);
    k_zero<<<grid1(n), 256, 0, st>>>(v, n);
    std::vector<int> perm(nsamples);
    std::mt19937_64 rng(seed);
    long step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (long i = 0; i < nsamples; ++i) perm[i] = (int)i;
        for (long i = nsamples - 1; i > 0; --i) {
            long j = (long)(rng() % (uint64_t)(i + 1));
            int tmp = perm[i]; perm[i] = perm[j]; perm[j] = tmp;
        }
        TDX_HIP(hipMemcpyAsync(idx, perm.data(), sizeof(int) * nsamples, hipMemcpyHostToDevice, st));
        for (long b0 = 0; b0 + batch <= nsamples; b0 += batch) {
            k_gather_rows<<<grid1(batch * nin), 256, 0, st>>>(X, idx + b0, xb, batch, nin);
            k_gather_rows<<<grid1(batch * nout), 256, 0, st>>>(T, idx + b0, tb, batch, nout);
            rc = backward(params, s, xb, tb, batch, w, G, loss, st);
            if (rc) return rc;
            ++step;
            const float c1 = (float)(1.0 - pow(0.9, (double)step));
            const float c2 = (float)(1.0 - pow(0.999, (double)step));
            k_adam<<<grid1(n), 256, 0, st>>>(params, G, m, v, n, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
        }
        if (loss_hist) {
            TDX_HIP(hipMemcpyAsync(&loss_hist[ep], loss, sizeof(float), hipMemcpyDeviceToHost, st));
        }
        // the host permutation buffer is rewritten next epoch: wait for this epoch's copies
        TDX_HIP(hipStreamSynchronize(st));
    }
    TDX_LAUNCH_CHECK();
    return record(TDX_OK);
}

}  // extern "C"
