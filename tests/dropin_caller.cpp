// A C++ caller written against the reference's surface (inc/cudecompose.h:23-52), in the
// style of its test programs (test/test_t2m.cpp, test/test_mttkrp.cpp, test/test_ttm.cpp):
// the caller allocates device buffers itself, fills them, calls the library functions,
// and copies the results back.  Test infrastructure: tests/test_gpu_dropin.py runs it and
// compares every output with the oracle.
//
// Usage: dropin_caller <out.bin>.  Inputs are small integers, so every product and sum is
// exact in FP32.  The file holds the inputs, then the outputs, as raw float32 arrays in the
// order written below, then the rse error.  Exit status 0 on success; otherwise the
// failing call is named on stderr.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cudecompose.h"

static const long M = 5, N = 4, K = 3, R = 3, D = 2;

static std::vector<dt> ints(long n, long mul, long add) {
    std::vector<dt> v(n);
    for (long i = 0; i < n; ++i) v[i] = (dt)(((i * mul + add) % 9) - 4);
    return v;
}

static dt *dev(const std::vector<dt> &h) {
    dt *p = nullptr;
    if (hipMalloc((void **)&p, h.size() * sizeof(dt)) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        exit(3);
    }
    if (hipMemcpy(p, h.data(), h.size() * sizeof(dt), hipMemcpyHostToDevice) != hipSuccess) {
        fprintf(stderr, "hipMemcpy failed\n");
        exit(3);
    }
    return p;
}

static FILE *out_file;

static void put(const std::vector<dt> &h) { fwrite(h.data(), sizeof(dt), h.size(), out_file); }

// Copies n floats of a device result back, checks the call's status and appends them.
static void take(const char *what, dt *d, long n) {
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "%s: status %d\n", what, tdx_last_status());
        exit(4);
    }
    std::vector<dt> h(n);
    if (hipMemcpy(h.data(), d, n * sizeof(dt), hipMemcpyDeviceToHost) != hipSuccess) {
        fprintf(stderr, "%s: copy back failed\n", what);
        exit(3);
    }
    put(h);
}

int main(int argc, char **argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: dropin_caller <out.bin>\n");
        return 2;
    }
    out_file = fopen(argv[1], "wb");
    if (!out_file) return 2;
    const long dims[3] = {M, N, K};
    std::vector<dt> X = ints(M * N * K, 7, 3), A = ints(M * R, 5, 1), A2 = ints(M * R, 4, 2),
                    A3 = ints(M * R, 2, 5), B = ints(N * R, 3, 4), C = ints(K * R, 8, 6);
    std::vector<std::vector<dt>> U = {ints(D * M, 5, 2), ints(D * N, 7, 1), ints(D * K, 4, 3)};
    for (const auto *h : {&X, &A, &A2, &A3, &B, &C, &U[0], &U[1], &U[2]}) put(*h);

    dt *dX = dev(X), *dA = dev(A), *dA2 = dev(A2), *dA3 = dev(A3), *dB = dev(B), *dC = dev(C);
    dt *dU[3] = {dev(U[0]), dev(U[1]), dev(U[2])};
    dt *dout = dev(std::vector<dt>(M * N * K * R * R, 0.f));

    for (long flag = 1; flag <= 3; ++flag) {
        t2m(dX, dout, flag, M, N, K);
        take("t2m", dout, M * N * K);
    }
    kr(dA, dB, dout, M, N, R);
    take("kr", dout, M * N * R);
    hadmard(dA, dA2, dout, M, R);
    take("hadmard", dout, M * R);
    hadmard3(dA, dA2, dA3, dout, M, R);
    take("hadmard3", dout, M * R);
    tensor_product(dA, dB, dout, M, R, N, R);
    take("tensor_product", dout, M * R * N * R);
    dt *left[3] = {dC, dC, dB}, *right[3] = {dB, dA, dA};
    for (long flag = 1; flag <= 3; ++flag) {
        mttkrp(dX, left[flag - 1], right[flag - 1], dout, flag, M, N, K, R);
        take("mttkrp", dout, dims[flag - 1] * R);
    }
    for (long flag = 1; flag <= 3; ++flag) {
        ttm(dX, dU[flag - 1], dout, flag, M, N, K, D);
        take("ttm", dout, M * N * K / dims[flag - 1] * D);
    }
    // rse against a copy of X with one entry changed by 2: error = 2 / ||X||
    std::vector<dt> X1 = X;
    X1[7] += 2.f;
    dt *dX1 = dev(X1), err = -1.f;
    rse(dX, dX1, M * N * K, &err);
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "rse: status %d\n", tdx_last_status());
        return 4;
    }
    fwrite(&err, sizeof(dt), 1, out_file);
    fclose(out_file);
    for (dt *p : {dX, dA, dA2, dA3, dB, dC, dU[0], dU[1], dU[2], dout, dX1}) hipFree(p);
    printf("dropin-caller-ok\n");
    return 0;
}
