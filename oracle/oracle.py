"""Python binding of the FP64 CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never calls it.

All arrays are numpy, column-major flattened (Fortran order), matching the
reference's storage convention (src/tensor2matrix.cu:4-36, src/cp_als.cu:5-6).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_L = ctypes.c_long


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_uniform.restype = ctypes.c_float
        L.orc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
        L.orc_normal.restype = ctypes.c_double
        L.orc_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
        L.orc_fill_uniform.argtypes = [_f32p, _L, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_fill_normal.argtypes = [_f32p, _L, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.orc_t2m.argtypes = [_f32p, _f32p, _L, _L, _L, _L]
        L.orc_kr.argtypes = [_f32p, _f32p, _f32p, _L, _L, _L]
        L.orc_hadamard.argtypes = [_f32p, _f32p, _f32p, _L]
        L.orc_hadamard3.argtypes = [_f32p, _f32p, _f32p, _f32p, _L]
        L.orc_tensor_product.argtypes = [_f32p, _f32p, _f32p, _L, _L, _L, _L]
        L.orc_tminust.argtypes = [_f32p, _f32p, _L]
        L.orc_identity.argtypes = [_f32p, _L]
        L.orc_f32_to_bf16.restype = ctypes.c_uint16
        L.orc_f32_to_bf16.argtypes = [ctypes.c_float]
        L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f32_to_f16.argtypes = [ctypes.c_float]
        L.orc_f32_to_e4m3.restype = ctypes.c_uint8
        L.orc_f32_to_e4m3.argtypes = [ctypes.c_float]
        L.orc_e4m3_to_f32.restype = ctypes.c_float
        L.orc_e4m3_to_f32.argtypes = [ctypes.c_uint8]
        L.orc_quantize_e4m3_mode1.argtypes = [_f32p, _L, _L, _u8p, _f32p]
        PP = ctypes.POINTER(ctypes.c_void_p)
        L.orc_mttkrp.argtypes = [_f32p, ctypes.c_int, ctypes.POINTER(_L), PP, _L, ctypes.c_int, _f64p]
        L.orc_mttkrp_scaled.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.POINTER(_L), PP, _L,
                                        ctypes.c_int, _f64p]
        L.orc_ttm.argtypes = [_f64p, _f64p, _f64p, _L, _L, _L, _L, _L]
        L.orc_gram.argtypes = [_f64p, _L, _L, _f64p]
        L.orc_cholesky.restype = ctypes.c_int
        L.orc_cholesky.argtypes = [_f64p, _L]
        L.orc_chol_solve_rows.argtypes = [_f64p, _L, _f64p, _L, _f64p]
        L.orc_pinv_solve_rows.argtypes = [_f64p, _L, _f64p, _L, _f64p, ctypes.c_double]
        L.orc_sym_eig.argtypes = [_f64p, _L, _f64p, _f64p]
        L.orc_cp_reconstruct.argtypes = [ctypes.c_int, ctypes.POINTER(_L), PP, _L, _f64p]
        L.orc_cp_relerr.restype = ctypes.c_double
        L.orc_cp_relerr.argtypes = [_f32p, ctypes.c_int, ctypes.POINTER(_L), PP, _L]
        L.orc_cp_als.restype = ctypes.c_int
        L.orc_cp_als.argtypes = [_f32p, ctypes.c_int, ctypes.POINTER(_L), _L, PP, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_void_p]
        L.orc_tucker_hosvd.argtypes = [_f32p, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_L), PP, _f64p]
        L.orc_tucker_hooi.argtypes = [_f32p, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_L),
                                      ctypes.c_int, PP, _f64p, ctypes.c_void_p]
        L.orc_tucker_reconstruct.argtypes = [ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_L), _f64p, PP, _f64p]
        L.orc_mlp_forward.argtypes = [_f64p, _L, _L, _L, _L] + [_f64p] * 7
        L.orc_mlp_backward.restype = ctypes.c_double
        L.orc_mlp_backward.argtypes = [_f64p, _f64p, _L, _L, _L, _L] + [_f64p] * 12
        L.orc_adam.argtypes = [_f64p, _f64p, _f64p, _f64p, _L, _L, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double]
        _LIB = L
    return _LIB


def _dims(dims):
    return (_L * 4)(*(list(dims) + [0] * (4 - len(dims))))


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def c32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel(order="F") if np.ndim(a) > 1 else
                                np.asarray(a, dtype=np.float32))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def uniform(n, seed, stream):
    out = np.empty(n, np.float32)
    lib().orc_fill_uniform(out, n, seed, stream)
    return out


def normal(n, seed, stream):
    out = np.empty(n, np.float32)
    lib().orc_fill_normal(out, n, seed, stream)
    return out


def t2m(X, flag, m, n, k):
    out = np.empty(m * n * k, np.float32)
    lib().orc_t2m(np.ascontiguousarray(X, np.float32), out, flag, m, n, k)
    return out


def kr(A, B, a, b, r):
    out = np.empty(a * b * r, np.float32)
    lib().orc_kr(np.ascontiguousarray(A, np.float32), np.ascontiguousarray(B, np.float32), out, a, b, r)
    return out


def hadamard(A, B):
    out = np.empty(A.size, np.float32)
    lib().orc_hadamard(np.ascontiguousarray(A, np.float32), np.ascontiguousarray(B, np.float32), out, A.size)
    return out


def hadamard3(A, B, C):
    out = np.empty(A.size, np.float32)
    lib().orc_hadamard3(*[np.ascontiguousarray(x, np.float32) for x in (A, B, C)], out, A.size)
    return out


def tensor_product(A, B, a, b, c, d):
    out = np.empty(a * b * c * d, np.float32)
    lib().orc_tensor_product(np.ascontiguousarray(A, np.float32), np.ascontiguousarray(B, np.float32),
                             out, a, b, c, d)
    return out


def tminust(L, R):
    out = np.ascontiguousarray(R, np.float32).copy()
    lib().orc_tminust(np.ascontiguousarray(L, np.float32), out, out.size)
    return out


def identity(m):
    out = np.empty(m * m, np.float32)
    lib().orc_identity(out, m)
    return out


def bf16_bits(x):
    f = lib().orc_f32_to_bf16
    return np.array([f(float(v)) for v in np.asarray(x, np.float32).ravel()], np.uint16)


def f16_bits(x):
    f = lib().orc_f32_to_f16
    return np.array([f(float(v)) for v in np.asarray(x, np.float32).ravel()], np.uint16)


def e4m3_bits(x):
    f = lib().orc_f32_to_e4m3
    return np.array([f(float(v)) for v in np.asarray(x, np.float32).ravel()], np.uint8)


def e4m3_decode(q):
    f = lib().orc_e4m3_to_f32
    return np.array([f(int(v)) for v in np.asarray(q, np.uint8).ravel()], np.float32)


def quantize_e4m3_mode1(X, I):
    X = np.ascontiguousarray(X, np.float32)
    Q = np.empty(X.size, np.uint8)
    s = np.empty(I, np.float32)
    lib().orc_quantize_e4m3_mode1(X, I, X.size // I, Q, s)
    return Q, s


def mttkrp(X, dims, U, mode, scales=None):
    """FP64 MTTKRP; X float32 flat col-major, U list of (I_m, R) arrays. Returns (I_mode, R) float64."""
    R = U[0].shape[1]
    Ud = [np.ascontiguousarray(np.asarray(u, np.float64).ravel(order="F")) for u in U]
    out = np.empty(dims[mode] * R, np.float64)
    X = np.ascontiguousarray(X, np.float32)
    if scales is None:
        lib().orc_mttkrp(X, len(dims), _dims(dims), _ptrs(Ud), R, mode, out)
    else:
        lib().orc_mttkrp_scaled(X, np.ascontiguousarray(scales, np.float32), len(dims), _dims(dims),
                                _ptrs(Ud), R, mode, out)
    return out.reshape((dims[mode], R), order="F")


def ttm(X, U, flag, a, b, c, d):
    osz = {1: d * b * c, 2: a * d * c, 3: a * b * d}[flag]
    out = np.empty(osz, np.float64)
    lib().orc_ttm(np.ascontiguousarray(X, np.float64), np.ascontiguousarray(U, np.float64), out, flag, a, b, c, d)
    return out


def sym_eig(A):
    n = A.shape[0]
    w = np.empty(n)
    V = np.empty(n * n)
    lib().orc_sym_eig(np.ascontiguousarray(np.asarray(A, np.float64).ravel(order="F")), n, w, V)
    return w, V.reshape((n, n), order="F")


def cholesky_solve_rows(V, M):
    R = V.shape[0]
    I = M.shape[0]
    L = np.ascontiguousarray(np.asarray(V, np.float64).ravel(order="F")).copy()
    st = lib().orc_cholesky(L, R)
    U = np.empty(I * R)
    if st == 0:
        lib().orc_chol_solve_rows(L, R, np.ascontiguousarray(M.ravel(order="F"), np.float64), I, U)
    else:
        lib().orc_pinv_solve_rows(np.ascontiguousarray(V.ravel(order="F"), np.float64), R,
                                  np.ascontiguousarray(M.ravel(order="F"), np.float64), I, U, 1e-12)
    return st, U.reshape((I, R), order="F")


def cp_reconstruct(dims, U):
    R = U[0].shape[1]
    Ud = [np.ascontiguousarray(np.asarray(u, np.float64).ravel(order="F")) for u in U]
    out = np.empty(int(np.prod(dims)), np.float64)
    lib().orc_cp_reconstruct(len(dims), _dims(dims), _ptrs(Ud), R, out)
    return out


def cp_relerr(X, dims, U):
    R = U[0].shape[1]
    Ud = [np.ascontiguousarray(np.asarray(u, np.float64).ravel(order="F")) for u in U]
    return lib().orc_cp_relerr(np.ascontiguousarray(X, np.float32), len(dims), _dims(dims), _ptrs(Ud), R)


def cp_als(X, dims, U0, sweeps, direct=False):
    """FP64 ALS from init U0 (list of (I_m,R)); returns (U list, fit relerr per sweep, direct relerr or None, status)."""
    R = U0[0].shape[1]
    Ud = [np.ascontiguousarray(np.asarray(u, np.float64).ravel(order="F")).copy() for u in U0]
    fit = np.empty(sweeps)
    dr = np.empty(sweeps) if direct else None
    st = lib().orc_cp_als(np.ascontiguousarray(X, np.float32), len(dims), _dims(dims), R, _ptrs(Ud), sweeps,
                          fit.ctypes.data, dr.ctypes.data if direct else None)
    U = [u.reshape((d, R), order="F") for u, d in zip(Ud, dims)]
    return U, fit, dr, st


def tucker_hosvd(X, dims, ranks):
    Ud = [np.empty(d * r) for d, r in zip(dims, ranks)]
    core = np.empty(int(np.prod(ranks)))
    lib().orc_tucker_hosvd(np.ascontiguousarray(X, np.float32), len(dims), _dims(dims), _dims(ranks),
                           _ptrs(Ud), core)
    return [u.reshape((d, r), order="F") for u, d, r in zip(Ud, dims, ranks)], core


def tucker_hooi(X, dims, ranks, iters, U0):
    Ud = [np.ascontiguousarray(np.asarray(u, np.float64).ravel(order="F")).copy() for u in U0]
    core = np.empty(int(np.prod(ranks)))
    hist = np.empty(max(iters, 1))
    lib().orc_tucker_hooi(np.ascontiguousarray(X, np.float32), len(dims), _dims(dims), _dims(ranks), iters,
                          _ptrs(Ud), core, hist.ctypes.data)
    return [u.reshape((d, r), order="F") for u, d, r in zip(Ud, dims, ranks)], core, hist[:iters]


def tucker_reconstruct(dims, ranks, core, U):
    Ud = [np.ascontiguousarray(np.asarray(u, np.float64).ravel(order="F")) for u in U]
    out = np.empty(int(np.prod(dims)))
    lib().orc_tucker_reconstruct(len(dims), _dims(dims), _dims(ranks), np.ascontiguousarray(core, np.float64),
                                 _ptrs(Ud), out)
    return out


def mlp_forward(x, params, nin, width, nout):
    W1, b1, W2, b2, W3, b3 = [np.ascontiguousarray(p, np.float64) for p in params]
    batch = x.shape[0]
    y = np.empty(batch * nout)
    lib().orc_mlp_forward(np.ascontiguousarray(x, np.float64).ravel(), batch, nin, width, nout,
                          W1, b1, W2, b2, W3, b3, y)
    return y.reshape(batch, nout)


def mlp_backward(x, t, params, nin, width, nout):
    W1, b1, W2, b2, W3, b3 = [np.ascontiguousarray(p, np.float64) for p in params]
    batch = x.shape[0]
    g = [np.empty_like(p) for p in (W1, b1, W2, b2, W3, b3)]
    loss = lib().orc_mlp_backward(np.ascontiguousarray(x, np.float64).ravel(),
                                  np.ascontiguousarray(t, np.float64).ravel(), batch, nin, width, nout,
                                  W1, b1, W2, b2, W3, b3, *g)
    return loss, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    lib().orc_adam(p, np.ascontiguousarray(g, np.float64), m, v, p.size, t, lr, b1, b2, eps)


def cp_als3_gemm(X, dims, U0, sweeps, dtype=np.float64, log=None, stop_at=None):
    """Order-3 CP-ALS with the MTTKRPs as GEMMs over the native layout: the formulation of the
    reference's CPU path, decomposition/CPU/python/mycp.py:11-22 (Hadamard of the other Grams,
    TensorLy's unfolding_dot_khatri_rao = X(n) (kr of the others), a linear solve per mode),
    with the reference library's mode order and conventions (src/cp_als.cu:44-119).

    Mode 0: M0 = X(0) (C (.) B), one GEMM.  Modes 1 and 2 contract A_new over i0 first:
    T = A_new^T X(0) (R x J K, one GEMM), then M1[j,r] = sum_k T[r,j,k] C[k,r] and
    M2[k,r] = sum_j T[r,j,k] B_new[j,r] -- the same sums as X(1) (C (.) A) and X(2) (B (.) A)
    in another order (exact algebra; FP64 rounding differences only).  The solve is the
    Cholesky solve of orc_cp_als (pseudo-inverse fallback).  Arithmetic in `dtype` (FP64: the
    oracle; FP32: the single-precision CPU baseline).  Returns (U list, per-sweep fast-fit
    relative error (Tensorlab cpd_als.m:281-283, accumulated in FP64), seconds per sweep).
    stop_at: stop after the first sweep whose relative error is <= stop_at."""
    import time
    I, J, K = (int(d) for d in dims)
    R = U0[0].shape[1]
    U = [np.array(u, dtype=np.float64, order="F") for u in U0]
    Xd = np.asarray(X).reshape(-1)
    if Xd.dtype != dtype:
        Xd = Xd.astype(dtype)
    X0 = Xd.reshape((I, J * K), order="F")           # mode-0 unfolding, a view
    normx2 = float(np.dot(Xd.astype(np.float64, copy=False), Xd.astype(np.float64, copy=False))) \
        if dtype == np.float64 else float(np.sum(np.square(Xd, dtype=np.float64)))

    def solve(V, M):
        try:
            L = np.linalg.cholesky(V)
            Y = np.linalg.solve(L, M.T)
            return np.linalg.solve(L.T, Y).T
        except np.linalg.LinAlgError:
            return (np.linalg.pinv(V, rcond=1e-12) @ M.T).T

    G = [u.T @ u for u in U]
    fit, secs = [], []
    for it in range(sweeps):
        t0 = time.perf_counter()
        A, B, C = U
        KR = (C.astype(dtype)[:, None, :] * B.astype(dtype)[None, :, :]).reshape(J * K, R)  # row j + k J
        M0 = np.asarray(X0 @ KR, np.float64)
        del KR
        U[0] = solve(G[1] * G[2], M0)
        G[0] = U[0].T @ U[0]
        T = (U[0].T.astype(dtype) @ X0).reshape(R, K, J)   # T[r, k, j]
        M1 = np.matmul(T.transpose(0, 2, 1), U[2].T.astype(dtype)[:, :, None])[:, :, 0].T  # (J, R)
        U[1] = solve(G[0] * G[2], np.asarray(M1, np.float64))
        G[1] = U[1].T @ U[1]
        M2 = np.matmul(T, U[1].T.astype(dtype)[:, :, None])[:, :, 0].T                  # (K, R)
        M2 = np.asarray(M2, np.float64)
        del T
        U[2] = solve(G[0] * G[1], M2)
        G[2] = U[2].T @ U[2]
        r2 = normx2 - 2.0 * float(np.sum(M2 * U[2])) + float(np.sum(G[0] * G[1] * G[2]))
        fit.append(float(np.sqrt(abs(r2) / normx2)))
        secs.append(time.perf_counter() - t0)
        if log:
            log(f"  sweep {it + 1}: relerr {fit[-1]:.12f} ({secs[-1]:.2f} s)")
        if stop_at is not None and fit[-1] <= stop_at:
            break
    return U, np.asarray(fit), np.asarray(secs)


def _unfold_gram_gemm(X, dims, mode):
    """G = X(mode) X(mode)^T in FP64 as one tensordot over the other indices."""
    P = int(np.prod(dims[:mode]))
    Q = int(np.prod(dims[mode + 1:]))
    Y = X.reshape((P, dims[mode], Q), order="F")
    if P == 1:
        Y2 = Y[0]                       # (In, Q)
        return Y2 @ Y2.T
    if Q == 1:
        Y2 = Y[:, :, 0]                 # (P, In)
        return Y2.T @ Y2
    return np.tensordot(Y, Y, axes=([0, 2], [0, 2]))


def _top_eigvecs_lapack(G, r):
    """The last r eigenvector columns in ascending order with the largest-magnitude entry of
    each made positive: the convention of orc_sym_eig / top_eigvecs (src/tucker_hosvd.cu:68-70)."""
    w, V = np.linalg.eigh(G)
    V = V[:, G.shape[0] - r:].copy()
    k = np.argmax(np.abs(V), axis=0)
    V *= np.where(V[k, np.arange(r)] < 0, -1.0, 1.0)
    return V, w[G.shape[0] - r:]


def _ttm_t_gemm(X, dims, mode, U):
    """Y = X x_mode U^T (U is I_mode x r), F-order flat in, F-order flat out."""
    P = int(np.prod(dims[:mode]))
    Q = int(np.prod(dims[mode + 1:]))
    Y = X.reshape((P, dims[mode], Q), order="F")
    Z = np.einsum("piq,ir->prq", Y, U, optimize=True) if P > 1 else np.tensordot(U, Y[0], axes=([0], [0]))[None]
    return np.asarray(Z).reshape(-1, order="F")


def tucker_hosvd_gemm(X, dims, ranks):
    """FP64 HOSVD (src/tucker_hosvd.cu:3-103, src/tucker4_hosvd.cu:3-141) with the Grams and
    the TTM chain as GEMMs and the eigensolve by LAPACK (eigh): the same quantities as
    orc_tucker_hosvd (whose per-pair Gram loops do not scale to 128^4), same conventions;
    tests/test_oracle.py pins one against the other.  Returns (U list, core flat F-order,
    top eigenvalues per mode)."""
    Xd = np.asarray(X, np.float64).reshape(-1)
    U, W = [], []
    for m in range(len(dims)):
        V, w = _top_eigvecs_lapack(_unfold_gram_gemm(Xd, dims, m), ranks[m])
        U.append(V)
        W.append(w)
    cur = list(dims)
    Y = Xd
    for m in range(len(dims)):
        Y = _ttm_t_gemm(Y, cur, m, U[m])
        cur[m] = ranks[m]
    return U, Y, W

def tucker_hooi_gemm(X, dims, ranks, iters, U0):
    """FP64 HOOI in GEMM form (the iteration of orc_tucker_hooi,
    decomposition/GPU/normal/tucker/tucker_hooi.cu:103-148: for each mode in order,
    Y = X x_{m != n} U_m^T with the current factors, U_n = the top eigenvectors of
    Y(n) Y(n)^T; the relative error sqrt(1 - ||core||^2 / ||X||^2) after every iteration),
    with the TTM chains and Grams as GEMMs and the eigensolve by LAPACK; pinned against
    orc_tucker_hooi in tests/test_oracle.py.  Returns (U list, core flat F-order, relerr
    per iteration, seconds per iteration)."""
    import time
    Xd = np.asarray(X, np.float64).reshape(-1)
    nx2 = float(np.dot(Xd, Xd))
    U = [np.array(u, np.float64) for u in U0]
    hist, secs = [], []

    def chain(skip):
        cur = list(dims)
        Y = Xd
        for m in range(len(dims)):
            if m == skip:
                continue
            Y = _ttm_t_gemm(Y, cur, m, U[m])
            cur[m] = ranks[m]
        return Y, cur

    for _ in range(iters):
        t0 = time.perf_counter()
        for n in range(len(dims)):
            Y, cur = chain(n)
            U[n], _ = _top_eigvecs_lapack(_unfold_gram_gemm(Y, cur, n), ranks[n])
        core, _ = chain(-1)
        hist.append(float(np.sqrt(max(nx2 - float(np.dot(core, core)), 0.0) / nx2)))
        secs.append(time.perf_counter() - t0)
    core, _ = chain(-1)
    return U, core, np.asarray(hist), np.asarray(secs)


def cp_als_gemm(X, dims, U0, sweeps, stop_at=None):
    """FP64 CP-ALS of any order in GEMM form (the reference CPU path,
    decomposition/CPU/python/mycp.py:11-22: for each mode in order, V = Hadamard of the other
    Grams, M = X(n) (Khatri-Rao of the others), a Cholesky solve; the fast-fit relative error
    of Tensorlab cpd_als.m:281-283 after every sweep), for order 4 where cp_als3_gemm does not
    apply; pinned against orc_cp_als in tests/test_oracle.py.  M = X(n) KR is formed as a
    tensordot of X with the Khatri-Rao rows of the longer side of mode n (the modes after it,
    or before it), then a contraction with the other side's.
    Returns (U list, relerr per sweep, seconds per sweep)."""
    import time
    d = len(dims)
    R = U0[0].shape[1]
    U = [np.array(u, np.float64) for u in U0]
    Xd = np.asarray(X, np.float64).reshape(-1)
    nx2 = float(np.dot(Xd, Xd))
    G = [u.T @ u for u in U]

    def mttkrp(n):
        P = int(np.prod(dims[:n]))
        Q = int(np.prod(dims[n + 1:]))
        Y = Xd.reshape((P, dims[n], Q), order="F")
        # Khatri-Rao rows of the later modes (index q = i_{n+1} + i_{n+2} I_{n+1} + ...)
        Kq = np.ones((1, R))
        for m in range(n + 1, d):
            Kq = (U[m][:, None, :] * Kq[None, :, :]).reshape(-1, R)
        Kp = np.ones((1, R))
        for m in range(n):
            Kp = (U[m][:, None, :] * Kp[None, :, :]).reshape(-1, R)
        # contract the longer side first, so the intermediate is In x min(P, Q) x R
        if Q >= P:
            Z = np.tensordot(Y, Kq, axes=([2], [0]))    # (P, In, R)
            return np.einsum("pir,pr->ir", Z, Kp, optimize=True)
        Z = np.tensordot(Y, Kp, axes=([0], [0]))        # (In, Q, R)
        return np.einsum("iqr,qr->ir", Z, Kq, optimize=True)

    def solve(V, M):
        try:
            L = np.linalg.cholesky(V)
            return np.linalg.solve(L.T, np.linalg.solve(L, M.T)).T
        except np.linalg.LinAlgError:
            return (np.linalg.pinv(V, rcond=1e-12) @ M.T).T

    fit, secs = [], []
    for _ in range(sweeps):
        t0 = time.perf_counter()
        for n in range(d):
            V = np.ones((R, R))
            for m in range(d):
                if m != n:
                    V = V * G[m]
            M = mttkrp(n)
            U[n] = solve(V, M)
            G[n] = U[n].T @ U[n]
        H = np.ones((R, R))
        for m in range(d):
            H = H * G[m]
        r2 = nx2 - 2.0 * float(np.sum(M * U[d - 1])) + float(np.sum(H))
        fit.append(float(np.sqrt(abs(r2) / nx2)))
        secs.append(time.perf_counter() - t0)
        if stop_at is not None and fit[-1] <= stop_at:
            break
    return U, np.asarray(fit), np.asarray(secs)
