"""A C++ program written against the reference's surface (include/cudecompose.h, which
restates inc/cudecompose.h:23-52) runs on the GPU with its own hipMalloc'd buffers, and
every output matches the oracle bit for bit.  The program (tests/dropin_caller.cpp,
built by __graft_entry__.build()) writes its integer inputs and its outputs to one file.

Oracle functions used: t2m (src/tensor2matrix.cu:38-52), kr (src/kr.cu:3-11), hadmard
and hadmard3 (src/hadmard.cu:10-30), tensor_product (src/tensorproduct.cu), mttkrp
(src/mttkrp.cu:3-80), ttm (src/ttm.cu:4-47).  rse (src/rse.cu:4-30) is checked against
its closed form, 2 / ||X|| for one entry changed by 2.  The drivers cp_als
(src/cp_als.cu:3-134) and tucker_hosvd (src/tucker_hosvd.cu:3-103) are checked by exact
low-rank recovery through gencp / gentucker (src/rse.cu:17-51) and rse."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "first-project_amd", "_build", "dropin_caller")
M, N, K, R, D = 5, 4, 3, 3, 2


@pytest.mark.gpu
def test_cpp_caller_through_cudecompose_h(tmp_path):
    assert os.path.exists(EXE), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    out = tmp_path / "dropin.bin"
    print("dropin_caller", out, flush=True)
    sys.stdout.flush()
    r = subprocess.run([EXE, str(out)], capture_output=True, text=True, timeout=90)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "dropin-caller-ok" in r.stdout
    buf = np.fromfile(str(out), dtype=np.float32)
    pos = 0

    def take(n):
        nonlocal pos
        v = buf[pos:pos + n]
        pos += n
        assert v.size == n, "file too short"
        return v

    dims = (M, N, K)
    X, A, A2, A3, B, C = take(M * N * K), take(M * R), take(M * R), take(M * R), take(N * R), take(K * R)
    U = [take(D * d) for d in dims]
    assert np.array_equal(X, np.round(X)) and np.abs(X).max() <= 4  # integer inputs as generated
    for flag in (1, 2, 3):
        assert np.array_equal(take(M * N * K), O.t2m(X, flag, M, N, K)), ("t2m", flag)
    assert np.array_equal(take(M * N * R), O.kr(A, B, M, N, R))
    assert np.array_equal(take(M * R), O.hadamard(A, A2))
    assert np.array_equal(take(M * R), O.hadamard3(A, A2, A3))
    assert np.array_equal(take(M * R * N * R), O.tensor_product(A, B, M, R, N, R))
    F = [a.reshape((d, R), order="F").astype(np.float64) for a, d in zip((A, B, C), dims)]
    for flag in (1, 2, 3):
        ref = O.mttkrp(X, dims, F, flag - 1).ravel(order="F").astype(np.float32)
        assert np.array_equal(take(dims[flag - 1] * R), ref), ("mttkrp", flag)
    for flag in (1, 2, 3):
        ref = O.ttm(X, U[flag - 1], flag, M, N, K, D).astype(np.float32)
        assert np.array_equal(take(M * N * K // dims[flag - 1] * D), ref), ("ttm", flag)
    err = take(1)[0]
    assert abs(err - 2.0 / np.linalg.norm(X.astype(np.float64))) < 1e-6
    # exact rank-2 CP and multilinear-rank-(2,2,2) tensors (12 x 10 x 8, integer factors with
    # column congruence <= 0.46): cp_als and tucker_hosvd recover them, so the rebuilt
    # tensor's relative error is at FP32 rounding level
    e_cp, e_tucker = take(2)
    print("cp_als relerr", e_cp, "tucker_hosvd relerr", e_tucker)
    assert 0 <= e_cp < 1e-4
    assert 0 <= e_tucker < 1e-5
    # f2h (src/float2half.cu:14-18): IEEE binary16 RNE bit patterns
    Xf = take(M * N * K)
    assert np.array_equal(take(M * N * K).astype(np.uint16), O.f16_bits(Xf))
    # tminust (src/tminust.cu:10-15): right := left - right; initMat (src/tminust.cu:31-36)
    assert np.array_equal(take(M * R), O.tminust(A, A2))
    assert np.array_equal(take(36), O.identity(6))
    # gencp4 / gentucker4 (src/rse.cu:53-96) on integer factors and core: exact
    P4, Q4, S4, T4, R4 = 7, 6, 5, 4, 2
    d4 = (P4, Q4, S4, T4)
    F4 = [take(d * R4).reshape((d, R4), order="F").astype(np.float64) for d in d4]
    G4 = take(R4 ** 4).astype(np.float64)
    assert np.array_equal(take(P4 * Q4 * S4 * T4), O.cp_reconstruct(d4, F4).astype(np.float32))
    rec_t4 = take(P4 * Q4 * S4 * T4)
    assert np.array_equal(rec_t4, O.tucker_reconstruct(d4, (R4,) * 4, G4, F4).astype(np.float32))
    e_cp4, e_cptc, e_tucker4, e_tuckertc = take(4)
    assert pos == buf.size
    print("cp_als4", e_cp4, "cp_tensorcore", e_cptc, "tucker_hosvd4", e_tucker4, "tucker_tensorcore", e_tuckertc)
    # exact low-rank recovery: rank-2 order-4 CP (500 FP32 sweeps), rank-2 order-3 CP with BF16
    # storage (integer entries <= 2^8 are exact in BF16) and 2 FP32 refinement sweeps,
    # multilinear rank (2,2,2,2); tucker_tensorcore takes its subspaces from the BF16 copy
    # (entries up to ~2^11 round at 2^-9 relative), so its error is that rounding's first order
    assert 0 <= e_cp4 < 1e-4
    assert 0 <= e_cptc < 1e-4
    assert 0 <= e_tucker4 < 1e-5
    assert 0 <= e_tuckertc < 1e-2
