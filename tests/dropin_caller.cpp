// A C++ caller written against the reference's surface (inc/cudecompose.h:23-52), in the
// style of its test programs (test/test_t2m.cpp, test/test_mttkrp.cpp, test/test_ttm.cpp):
// the caller allocates device buffers itself, fills them, calls the library functions,
// and copies the results back.  Test infrastructure: tests/test_gpu_dropin.py runs it and
// compares every output with the oracle.
//
// Usage: dropin_caller <out.bin>.  Inputs are small integers, so every product and sum is
// exact in FP32 (f2h alone gets non-integers).  The file holds the inputs, then the outputs,
// as raw float32 arrays in the order written below, the rse error, the relative errors of
// the cp_als and tucker_hosvd reconstructions, then the f2h / tminust / initMat / gencp4 /
// gentucker4 inputs and outputs and the relative errors of the cp_als4, cp_tensorcore,
// tucker_hosvd4 and tucker_tensorcore reconstructions.  All 22 functions of the reference
// header run.  Exit status 0 on success; otherwise the failing call is named on stderr.
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

    // Drivers, as the reference's test/test_cp.cpp and test/test_tucker.cpp use them: build an
    // exact low-rank tensor, decompose it, rebuild it from the result and take rse.
    const long P = 12, Q = 10, S = 8, RC = 2, RT = 2;
    dt *dXc = dev(std::vector<dt>(P * Q * S, 0.f)), *drec = dev(std::vector<dt>(P * Q * S, 0.f));
    dt *dFa = dev(ints(P * RC, 5, 2)), *dFb = dev(ints(Q * RC, 7, 3)), *dFc = dev(ints(S * RC, 4, 1));
    dt *dOa = dev(std::vector<dt>(P * RC, 0.f)), *dOb = dev(std::vector<dt>(Q * RC, 0.f)),
       *dOc = dev(std::vector<dt>(S * RC, 0.f));
    dt errs[2] = {-1.f, -1.f};
    gencp(dXc, dFa, dFb, dFc, P, Q, S, RC);
    cp_als(dXc, dOa, dOb, dOc, P, Q, S, RC);
    gencp(drec, dOa, dOb, dOc, P, Q, S, RC);
    rse(dXc, drec, P * Q * S, &errs[0]);
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "cp_als chain: status %d\n", tdx_last_status());
        return 4;
    }
    dt *dG = dev(ints(RT * RT * RT, 5, 4)), *dT1 = dev(std::vector<dt>(P * RT, 0.f)),
       *dT2 = dev(std::vector<dt>(Q * RT, 0.f)), *dT3 = dev(std::vector<dt>(S * RT, 0.f)),
       *dG2 = dev(std::vector<dt>(RT * RT * RT, 0.f));
    gentucker(dXc, dG, dFa, dFb, dFc, P, Q, S, RT, RT, RT);
    tucker_hosvd(dXc, dG2, dT1, dT2, dT3, P, Q, S, RT, RT, RT);
    gentucker(drec, dG2, dT1, dT2, dT3, P, Q, S, RT, RT, RT);
    rse(dXc, drec, P * Q * S, &errs[1]);
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "tucker_hosvd chain: status %d\n", tdx_last_status());
        return 4;
    }
    fwrite(errs, sizeof(dt), 2, out_file);

    // The remaining entry points: f2h, tminust, initMat, printTensor, the order-4 builders
    // gencp4 / gentucker4 (outputs compared with the oracle), and the drivers cp_als4,
    // cp_tensorcore, tucker_hosvd4, tucker_tensorcore by low-rank recovery as above.
    half *dh = nullptr;
    if (hipMalloc((void **)&dh, M * N * K * sizeof(half)) != hipSuccess) return 3;
    std::vector<dt> Xf = X;
    for (long i = 0; i < M * N * K; ++i) Xf[i] = X[i] * 0.3f + 1e-3f * (dt)i;  // non-integers: rounding matters
    dt *dXf = dev(Xf);
    put(Xf);
    f2h(dXf, dh, M * N * K);
    {
        std::vector<half> hh(M * N * K);
        if (tdx_last_status() != TDX_OK || hipMemcpy(hh.data(), dh, hh.size() * sizeof(half), hipMemcpyDeviceToHost) != hipSuccess) {
            fprintf(stderr, "f2h: status %d\n", tdx_last_status());
            return 4;
        }
        std::vector<dt> hf(hh.begin(), hh.end());  // bit patterns as exact floats
        put(hf);
    }
    dt *dL = dev(A), *dRt = dev(A2);
    tminust(dL, dRt, M * R);
    take("tminust", dRt, M * R);
    initMat(dout, 6);
    take("initMat", dout, 36);
    printTensor(dA, 2, 2, 2);
    fflush(stdout);
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "printTensor: status %d\n", tdx_last_status());
        return 4;
    }

    const long P4 = 7, Q4 = 6, S4 = 5, T4 = 4, R4 = 2;
    std::vector<dt> Fa4 = ints(P4 * R4, 5, 2), Fb4 = ints(Q4 * R4, 7, 3), Fc4 = ints(S4 * R4, 4, 1),
                    Fd4 = ints(T4 * R4, 2, 7), G4 = ints(R4 * R4 * R4 * R4, 5, 4);
    for (const auto *h : {&Fa4, &Fb4, &Fc4, &Fd4, &G4}) put(*h);
    const long N4 = P4 * Q4 * S4 * T4;
    dt *dX4 = dev(std::vector<dt>(N4, 0.f)), *drec4 = dev(std::vector<dt>(N4, 0.f));
    dt *dFa4 = dev(Fa4), *dFb4 = dev(Fb4), *dFc4 = dev(Fc4), *dFd4 = dev(Fd4), *dG4 = dev(G4);
    dt *dO4[4] = {dev(std::vector<dt>(P4 * R4, 0.f)), dev(std::vector<dt>(Q4 * R4, 0.f)),
                  dev(std::vector<dt>(S4 * R4, 0.f)), dev(std::vector<dt>(T4 * R4, 0.f))};
    dt *dGo4 = dev(std::vector<dt>(R4 * R4 * R4 * R4, 0.f));
    dt errs2[4] = {-1.f, -1.f, -1.f, -1.f};
    gencp4(dX4, dFa4, dFb4, dFc4, dFd4, P4, Q4, S4, T4, R4);
    take("gencp4", dX4, N4);
    cp_als4(dX4, dO4[0], dO4[1], dO4[2], dO4[3], P4, Q4, S4, T4, R4);
    gencp4(drec4, dO4[0], dO4[1], dO4[2], dO4[3], P4, Q4, S4, T4, R4);
    rse(dX4, drec4, N4, &errs2[0]);
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "cp_als4 chain: status %d\n", tdx_last_status());
        return 4;
    }
    gencp(dXc, dFa, dFb, dFc, P, Q, S, RC);
    cp_tensorcore(dXc, dOa, dOb, dOc, P, Q, S, RC);
    gencp(drec, dOa, dOb, dOc, P, Q, S, RC);
    rse(dXc, drec, P * Q * S, &errs2[1]);
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "cp_tensorcore chain: status %d\n", tdx_last_status());
        return 4;
    }
    gentucker4(dX4, dG4, dFa4, dFb4, dFc4, dFd4, P4, Q4, S4, T4, R4, R4, R4, R4);
    take("gentucker4", dX4, N4);
    tucker_hosvd4(dX4, dGo4, dO4[0], dO4[1], dO4[2], dO4[3], P4, Q4, S4, T4, R4, R4, R4, R4);
    gentucker4(drec4, dGo4, dO4[0], dO4[1], dO4[2], dO4[3], P4, Q4, S4, T4, R4, R4, R4, R4);
    rse(dX4, drec4, N4, &errs2[2]);
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "tucker_hosvd4 chain: status %d\n", tdx_last_status());
        return 4;
    }
    gentucker(dXc, dG, dFa, dFb, dFc, P, Q, S, RT, RT, RT);
    tucker_tensorcore(dXc, dG2, dT1, dT2, dT3, P, Q, S, RT, RT, RT);
    gentucker(drec, dG2, dT1, dT2, dT3, P, Q, S, RT, RT, RT);
    rse(dXc, drec, P * Q * S, &errs2[3]);
    if (tdx_last_status() != TDX_OK) {
        fprintf(stderr, "tucker_tensorcore chain: status %d\n", tdx_last_status());
        return 4;
    }
    fwrite(errs2, sizeof(dt), 4, out_file);
    fclose(out_file);
    if (hipFree(dh) != hipSuccess) return 3;
    for (dt *p : {dX, dA, dA2, dA3, dB, dC, dU[0], dU[1], dU[2], dout, dX1, dXc, drec, dFa, dFb, dFc, dOa, dOb,
                  dOc, dG, dT1, dT2, dT3, dG2, dXf, dL, dRt, dX4, drec4, dFa4, dFb4, dFc4, dFd4, dG4, dO4[0],
                  dO4[1], dO4[2], dO4[3], dGo4})
        if (hipFree(p) != hipSuccess) return 3;
    printf("dropin-caller-ok\n");
    return 0;
}
