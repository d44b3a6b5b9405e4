"""ctypes binding of libtdx.so (the MI355X-native tensor-decomposition library).

The library takes raw device pointers; here torch tensors provide device memory and
the default HIP stream.  Every call raises TdxError on a non-zero status, and loading
fails loudly when libtdx.so has not been built: there is no CPU fallback.

Names mirror the reference surface (inc/cudecompose.h): hadmard, hadmard3, kr,
tensor_product, t2m, tminust, initMat, f2h, ttm, mttkrp, rse, gencp, gencp4,
gentucker, gentucker4, cp_als, cp_als4, cp_tensorcore, tucker_hosvd, tucker_hosvd4,
tucker_tensorcore.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtdx.so")
_LIB = None

TDX_F32, TDX_BF16, TDX_FP8_E4M3, TDX_F16 = 0, 1, 2, 3
(TDX_OK, TDX_ERR_ARG, TDX_ERR_ALLOC, TDX_ERR_HIP, TDX_WARN_SINGULAR, TDX_ERR_NONFINITE, TDX_ERR_UNSUPPORTED,
 TDX_ERR_TOO_LARGE, TDX_ERR_IO) = range(9)

_L = ctypes.c_long
_P = ctypes.c_void_p


class TdxError(RuntimeError):
    def __init__(self, status, fn):
        self.status = status
        msg = lib().tdx_status_string(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{fn}: status {status} ({msg})")


LS_FEATURES = 12
LS_NONE, LS_LEARNED, LS_LSB, LS_GRID = 0, 1, 2, 3


class TensorInfo(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("storage", ctypes.c_int), ("dims", ctypes.c_int64 * 4),
                ("n_scales", ctypes.c_int64), ("data_bytes", ctypes.c_int64)]


class CpTrace(ctypes.Structure):
    _fields_ = [("alpha", _P), ("accepted", _P), ("features", _P), ("grid_relerr", _P), ("plain_relerr", _P),
                ("elapsed_ms", _P)]


class CpOptions(ctypes.Structure):
    _fields_ = [("max_sweeps", ctypes.c_int), ("refine_sweeps", ctypes.c_int), ("storage", ctypes.c_int),
                ("init", ctypes.c_int), ("seed", ctypes.c_uint64), ("tol", ctypes.c_double),
                ("record_fit", ctypes.c_int), ("line_search", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("mlp_params", _P), ("mlp_width", ctypes.c_int), ("ngrid", ctypes.c_int),
                ("grid", ctypes.c_double * 8), ("trace", ctypes.POINTER(CpTrace)), ("target_relerr", ctypes.c_double),
                ("shared_pass", ctypes.c_int), ("opening_sweeps", ctypes.c_int),
                ("opening_split", ctypes.c_int)]


class CpReport(ctypes.Structure):
    _fields_ = [("sweeps_done", ctypes.c_int), ("status", ctypes.c_int), ("relerr_fit", ctypes.c_double),
                ("seconds", ctypes.c_double), ("setup_seconds", ctypes.c_double)]


_SIGS = {
    "tdx_init": [], "tdx_set_stream": [_P], "tdx_last_status": [], "tdx_synchronize": [],
    "tdx_release_workspace": [], "tdx_copy_to_host": [_P, _P, ctypes.c_size_t],
    "tdx_copy_to_device": [_P, _P, ctypes.c_size_t],
    "tdx_hadamard": [_P, _P, _P, _L, _L],
    "tdx_hadamard3": [_P, _P, _P, _P, _L, _L],
    "tdx_kr": [_P, _P, _P, _L, _L, _L],
    "tdx_tensor_product": [_P, _P, _P, _L, _L, _L, _L],
    "tdx_t2m": [_P, _P, _L, _L, _L, _L],
    "tdx_tminust": [_P, _P, _L],
    "tdx_init_identity": [_P, _L],
    "tdx_f2h": [_P, _P, _L],
    "tdx_quantize_bf16": [_P, _P, _L],
    "tdx_dequantize_bf16": [_P, _P, _L],
    "tdx_quantize_fp8_mode1": [_P, _P, _P, _L, _L],
    "tdx_dequantize_fp8_mode1": [_P, _P, _P, _L, _L],
    "tdx_mttkrp": [_P, _P, _P, _P, _L, _L, _L, _L, _L],
    "tdx_mttkrp_ex": [_P, ctypes.c_int, _P, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_P), _L,
                      ctypes.c_int, _P],
    "tdx_tensor_info_read": [ctypes.c_char_p, _P],
    "tdx_tensor_save": [ctypes.c_char_p, _P, ctypes.c_int, _P, ctypes.c_int, ctypes.POINTER(_L), ctypes.c_int],
    "tdx_tensor_load": [ctypes.c_char_p, _P, ctypes.c_size_t, _P, ctypes.c_size_t, ctypes.c_int, _P],
    "tdx_mttkrp_ex_noalloc": [_P, ctypes.c_int, _P, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_P), _L,
                      ctypes.c_int, _P],
    "tdx_bench_mttkrp": [_P, ctypes.c_int, _P, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_P), _L,
                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)],
    "tdx_mttkrp_dtree": [_P, ctypes.c_int, _P, ctypes.POINTER(_L), ctypes.POINTER(_P), _L, _P, _P],
    "tdx_bench_mttkrp_dtree": [_P, ctypes.c_int, _P, ctypes.POINTER(_L), ctypes.POINTER(_P), _L, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "tdx_ttm": [_P, _P, _P, _L, _L, _L, _L, _L],
    "tdx_ttm_t": [_P, ctypes.c_int, ctypes.POINTER(_L), ctypes.c_int, _P, _L, _P],
    "tdx_unfold_gram": [_P, ctypes.c_int, ctypes.POINTER(_L), ctypes.c_int, _P],
    "tdx_sym_eig_topr": [_P, _L, _L, _P, _P],
    "tdx_sym_eig_topr_ex": [_P, _L, _L, _P, ctypes.c_int, _P, _P, _P],
    "tdx_ttm_ex": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_L), ctypes.c_int, _P, _L, _P],
    "tdx_unfold_gram_ex": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_L), ctypes.c_int, _P],
    "tdx_rse": [_P, _P, _L, ctypes.POINTER(ctypes.c_float)],
    "tdx_gencp": [_P, _P, _P, _P, _L, _L, _L, _L],
    "tdx_gencp4": [_P, _P, _P, _P, _P, _L, _L, _L, _L, _L],
    "tdx_gentucker": [_P, _P, _P, _P, _P, _L, _L, _L, _L, _L, _L],
    "tdx_gentucker4": [_P, _P, _P, _P, _P, _P, _L, _L, _L, _L, _L, _L, _L, _L],
    "tdx_cp_relerr": [_P, ctypes.c_int, _P, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_P), _L,
                      ctypes.POINTER(ctypes.c_double)],
    "tdx_fill_uniform": [_P, _L, ctypes.c_uint64, ctypes.c_uint32],
    "tdx_fill_normal": [_P, _L, ctypes.c_uint64, ctypes.c_uint32],
    "tdx_gen_cp_tensor": [_P, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_P), _L, ctypes.c_double,
                          ctypes.c_uint64, ctypes.c_uint32],
    "tdx_cp_als_ex": [_P, ctypes.c_int, ctypes.POINTER(_L), _L, ctypes.POINTER(_P), ctypes.POINTER(CpOptions),
                      ctypes.POINTER(CpReport), _P, _P],
    "tdx_cp_als": [_P, _P, _P, _P, _L, _L, _L, _L],
    "tdx_cp_als4": [_P, _P, _P, _P, _P, _L, _L, _L, _L, _L],
    "tdx_cp_tensorcore": [_P, _P, _P, _P, _L, _L, _L, _L],
    "tdx_tucker_hosvd": [_P, _P, _P, _P, _P, _L, _L, _L, _L, _L, _L],
    "tdx_tucker_hosvd4": [_P, _P, _P, _P, _P, _P, _L, _L, _L, _L, _L, _L, _L, _L],
    "tdx_tucker_tensorcore": [_P, _P, _P, _P, _P, _L, _L, _L, _L, _L, _L],
    "tdx_tucker_hooi": [_P, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_L), ctypes.c_int, ctypes.c_int,
                        _P, ctypes.POINTER(_P), _P],
    "tdx_cp_als_batched": [_P, _P, _P, _P, _P, _P, _L, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                           _P, _P],
    "tdx_cp_batched_autotune": [_P, _P, _P, _P, _P, _P, _L, _L, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                _P],
    "tdx_mlp_init": [_P, _L, _L, _L, ctypes.c_uint64],
    "tdx_mlp_forward": [_P, _P, _L, _L, _L, _L, _P],
    "tdx_mlp_backward": [_P, _P, _P, _L, _L, _L, _L, _P, _P],
    "tdx_mlp_adam": [_P, _P, _P, _P, _L, _L, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float],
    "tdx_mlp_train": [_P, _P, _P, _L, _L, _L, _L, ctypes.c_int, _L, ctypes.c_float, ctypes.c_uint64, _P],
}

# every symbol include/tdx.h declares (checked against the built library by the tests)
HEADER_SYMBOLS = sorted(list(_SIGS) + ["tdx_get_stream", "tdx_status_string", "tdx_cp_default_options",
                                        "tdx_mlp_param_count", "tdx_cp_batched_variants",
                                        "tdx_cp_batched_variant_name"])


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libtdx.so not built at {LIB_PATH}; run __graft_entry__.build()")
        # torch ships its own HIP runtime under the same soname as ROCm's; whichever is loaded
        # first serves the whole process, so load torch first and let libtdx bind to it (the
        # tensors this module passes are torch allocations)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.tdx_status_string.restype = ctypes.c_char_p
        L.tdx_status_string.argtypes = [ctypes.c_int]
        L.tdx_get_stream.restype = _P
        L.tdx_cp_default_options.argtypes = [ctypes.POINTER(CpOptions)]
        L.tdx_cp_default_options.restype = None
        L.tdx_mlp_param_count.restype = _L
        L.tdx_mlp_param_count.argtypes = [_L, _L, _L]
        L.tdx_cp_batched_variants.restype = ctypes.c_int
        L.tdx_cp_batched_variants.argtypes = []
        L.tdx_cp_batched_variant_name.restype = ctypes.c_char_p
        L.tdx_cp_batched_variant_name.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _call(name, *args):
    st = getattr(lib(), name)(*args)
    if st not in (TDX_OK,):
        raise TdxError(st, name)
    return st


def _call_status(name, *args):
    return getattr(lib(), name)(*args)


def ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def dims_arr(dims):
    return (_L * 4)(*(list(dims) + [1] * (4 - len(dims))))


def ptr_arr(ts):
    return (_P * 4)(*([t.data_ptr() if t is not None else None for t in ts] + [None] * (4 - len(ts))))


# ---------------------------------------------------------------------------
# reference-named wrappers on torch CUDA tensors (flat, column-major contents)

def hadmard(A, B, C, a, b):
    _call("tdx_hadamard", ptr(A), ptr(B), ptr(C), a, b)


def hadmard3(A, B, C, ABC, a, b):
    _call("tdx_hadamard3", ptr(A), ptr(B), ptr(C), ptr(ABC), a, b)


def kr(A, B, AkrB, a, b, r):
    _call("tdx_kr", ptr(A), ptr(B), ptr(AkrB), a, b, r)


def tensor_product(A, B, C, a, b, c, d):
    _call("tdx_tensor_product", ptr(A), ptr(B), ptr(C), a, b, c, d)


def t2m(T1, T2, flag, m, n, k):
    _call("tdx_t2m", ptr(T1), ptr(T2), flag, m, n, k)


def tminust(left, right, m):
    _call("tdx_tminust", ptr(left), ptr(right), m)


def initMat(A, m):
    _call("tdx_init_identity", ptr(A), m)


def f2h(A, B, num):
    _call("tdx_f2h", ptr(A), ptr(B), num)


def quantize_bf16(X, Xq):
    _call("tdx_quantize_bf16", ptr(X), ptr(Xq), X.numel())


def quantize_fp8_mode1(X, Xq, scales, I):
    _call("tdx_quantize_fp8_mode1", ptr(X), ptr(Xq), ptr(scales), I, X.numel() // I)


def dequantize_fp8_mode1(Xq, scales, X, I):
    _call("tdx_dequantize_fp8_mode1", ptr(Xq), ptr(scales), ptr(X), I, X.numel() // I)


def mttkrp(X, left, right, XU, flag, m, n, k, r):
    _call("tdx_mttkrp", ptr(X), ptr(left), ptr(right), ptr(XU), flag, m, n, k, r)


def mttkrp_ex(X, dtype, scales, dims, factors, rank, mode, out):
    _call("tdx_mttkrp_ex", ptr(X), dtype, ptr(scales), len(dims), dims_arr(dims), ptr_arr(factors), rank, mode,
          ptr(out))


def tensor_info(path):
    info = TensorInfo()
    _call("tdx_tensor_info_read", os.fsencode(path), ctypes.byref(info))
    return info


def _addr(a):
    """device tensor -> (pointer, on_device=1); numpy array -> (pointer, 0)"""
    if a is None:
        return None, 0
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr()), 1
    return ctypes.c_void_p(a.ctypes.data), 0


def tensor_save(path, X, storage, dims, scales=None):
    """Writes X (torch device tensor or numpy array, in `storage`) as a .tdxq file."""
    px, dev_ = _addr(X)
    ps, _ = _addr(scales)
    _call("tdx_tensor_save", os.fsencode(path), px, storage, ps, len(dims), dims_arr(dims), dev_)


def tensor_load(path, X, scales=None):
    """Loads a .tdxq file into X (and scales for FP8); both torch device tensors or both numpy
    arrays, sized in bytes at least data_bytes / 4 n_scales.  Returns the TensorInfo."""
    px, dev_ = _addr(X)
    ps, _ = _addr(scales)
    nx = X.numel() * X.element_size() if hasattr(X, "numel") else X.nbytes
    ns = 0 if scales is None else (scales.numel() * scales.element_size() if hasattr(scales, "numel")
                                   else scales.nbytes)
    info = TensorInfo()
    _call("tdx_tensor_load", os.fsencode(path), px, nx, ps, ns, dev_, ctypes.byref(info))
    return info


def mttkrp_ex_noalloc(X, dtype, scales, dims, factors, rank, mode, out):
    """tdx_mttkrp_ex with no device allocation (fixed 64 MiB scratch per GPU); returns the status
    (TDX_ERR_TOO_LARGE when the panels and partials do not fit)."""
    return _call_status("tdx_mttkrp_ex_noalloc", ptr(X), dtype, ptr(scales), len(dims), dims_arr(dims),
                        ptr_arr(factors), rank, mode, ptr(out))


def mttkrp_dtree(X, dtype, scales, dims, factors, rank, M1, M2):
    """Order 3: M1 (mode 1) and M2 (mode 2) from the shared pass (T = X x_0 A^T)."""
    _call("tdx_mttkrp_dtree", ptr(X), dtype, ptr(scales), dims_arr(dims), ptr_arr(factors), rank, ptr(M1), ptr(M2))


def bench_mttkrp_dtree(X, dtype, scales, dims, factors, rank, iters):
    """Mean launch ms of the shared pass's two kernels: (pass over X, pass over T)."""
    m1, m2 = ctypes.c_float(), ctypes.c_float()
    _call("tdx_bench_mttkrp_dtree", ptr(X), dtype, ptr(scales), dims_arr(dims), ptr_arr(factors), rank, iters,
          ctypes.byref(m1), ctypes.byref(m2))
    return m1.value, m2.value


def bench_mttkrp(X, dtype, scales, dims, factors, rank, mode, iters):
    ms = ctypes.c_float()
    _call("tdx_bench_mttkrp", ptr(X), dtype, ptr(scales), len(dims), dims_arr(dims), ptr_arr(factors), rank, mode,
          iters, ctypes.byref(ms))
    return ms.value


def ttm(X, U, XU, flag, a, b, c, d):
    _call("tdx_ttm", ptr(X), ptr(U), ptr(XU), flag, a, b, c, d)


def rse(X, X1, m):
    err = ctypes.c_float()
    _call("tdx_rse", ptr(X), ptr(X1), m, ctypes.byref(err))
    return err.value


def gencp(rec, A, B, C, a, b, c, r):
    _call("tdx_gencp", ptr(rec), ptr(A), ptr(B), ptr(C), a, b, c, r)


def gencp4(rec, A, B, C, D, a, b, c, d, r):
    _call("tdx_gencp4", ptr(rec), ptr(A), ptr(B), ptr(C), ptr(D), a, b, c, d, r)


def gentucker(rec, core, A, B, C, a, b, c, r1, r2, r3):
    _call("tdx_gentucker", ptr(rec), ptr(core), ptr(A), ptr(B), ptr(C), a, b, c, r1, r2, r3)


def gentucker4(rec, core, A, B, C, D, a, b, c, d, r1, r2, r3, r4):
    _call("tdx_gentucker4", ptr(rec), ptr(core), ptr(A), ptr(B), ptr(C), ptr(D), a, b, c, d, r1, r2, r3, r4)


def cp_relerr(X, dims, factors, rank, dtype=TDX_F32, scales=None):
    out = ctypes.c_double()
    _call("tdx_cp_relerr", ptr(X), dtype, ptr(scales), len(dims), dims_arr(dims), ptr_arr(factors), rank,
          ctypes.byref(out))
    return out.value


def fill_uniform(out, seed, stream):
    _call("tdx_fill_uniform", ptr(out), out.numel(), seed, stream)


def fill_normal(out, seed, stream):
    _call("tdx_fill_normal", ptr(out), out.numel(), seed, stream)


def gen_cp_tensor(X, dims, factors, rank, eta, seed, stream):
    _call("tdx_gen_cp_tensor", ptr(X), len(dims), dims_arr(dims), ptr_arr(factors), rank, float(eta), seed, stream)


def default_cp_options():
    o = CpOptions()
    lib().tdx_cp_default_options(ctypes.byref(o))
    return o


def cp_als_ex(X, dims, rank, factors, opts=None, want_fit=True, want_ms=False, trace=None, want_elapsed=False):
    """Returns (status, report, fit_hist, sweep_ms).  With trace=dict() and a line search
    the dict receives per-sweep arrays alpha, accepted, features, plain_relerr (and
    grid_relerr for the grid mode).  With want_elapsed, report.elapsed_ms holds the device ms
    from the call's first device work to the end of every sweep."""
    import numpy as np
    o = opts if opts is not None else default_cp_options()
    total = o.max_sweeps + (o.refine_sweeps if o.storage != TDX_F32 else 0)
    n = max(total, 1)
    fit = np.zeros(n, np.float64)
    ms = np.zeros(n, np.float32)
    elapsed = np.zeros(n, np.float32) if want_elapsed else None
    rep = CpReport()
    tr = None
    if (trace is not None and o.line_search) or want_elapsed:
        tr = CpTrace()
        if trace is not None and o.line_search:
            trace["alpha"] = np.zeros(n, np.float32)
            trace["accepted"] = np.zeros(n, np.int32)
            trace["features"] = np.zeros((n, LS_FEATURES), np.float32)
            trace["plain_relerr"] = np.zeros(n, np.float64)
            trace["grid_relerr"] = np.zeros((n, max(o.ngrid, 1)), np.float64)
            tr.alpha, tr.accepted = trace["alpha"].ctypes.data, trace["accepted"].ctypes.data
            tr.features, tr.grid_relerr = trace["features"].ctypes.data, trace["grid_relerr"].ctypes.data
            tr.plain_relerr = trace["plain_relerr"].ctypes.data
        if want_elapsed:
            tr.elapsed_ms = elapsed.ctypes.data
        o.trace = ctypes.pointer(tr)
    st = _call_status("tdx_cp_als_ex", ptr(X), len(dims), dims_arr(dims), rank, ptr_arr(factors), ctypes.byref(o),
                      ctypes.byref(rep), fit.ctypes.data if want_fit else None, ms.ctypes.data if want_ms else None)
    if tr is not None:
        o.trace = None
        if trace is not None and o.line_search:
            for k in trace:
                trace[k] = trace[k][:rep.sweeps_done]
    if want_elapsed:
        rep.elapsed_ms = elapsed[:rep.sweeps_done]
    if st not in (TDX_OK, TDX_WARN_SINGULAR):
        raise TdxError(st, "tdx_cp_als_ex")
    return st, rep, fit[:rep.sweeps_done], ms[:rep.sweeps_done]


def cp_als(X, A, B, C, a, b, c, r):
    return _call_status("tdx_cp_als", ptr(X), ptr(A), ptr(B), ptr(C), a, b, c, r)


def cp_als4(X, A, B, C, D, a, b, c, d, r):
    return _call_status("tdx_cp_als4", ptr(X), ptr(A), ptr(B), ptr(C), ptr(D), a, b, c, d, r)


def cp_tensorcore(X, A, B, C, a, b, c, r):
    return _call_status("tdx_cp_tensorcore", ptr(X), ptr(A), ptr(B), ptr(C), a, b, c, r)


def tucker_hosvd(X, core, U1, U2, U3, a, b, c, r1, r2, r3):
    _call("tdx_tucker_hosvd", ptr(X), ptr(core), ptr(U1), ptr(U2), ptr(U3), a, b, c, r1, r2, r3)


def tucker_hosvd4(X, core, U1, U2, U3, U4, a, b, c, d, r1, r2, r3, r4):
    _call("tdx_tucker_hosvd4", ptr(X), ptr(core), ptr(U1), ptr(U2), ptr(U3), ptr(U4), a, b, c, d, r1, r2, r3, r4)


def tucker_tensorcore(X, core, U1, U2, U3, a, b, c, r1, r2, r3):
    _call("tdx_tucker_tensorcore", ptr(X), ptr(core), ptr(U1), ptr(U2), ptr(U3), a, b, c, r1, r2, r3)


def sym_eig_topr(G, n, r, U, w=None):
    """U (n x r FP32) <- eigenvectors of the r largest eigenvalues of the FP64 n x n G,
    ascending; w (FP64, r) <- those eigenvalues."""
    _call("tdx_sym_eig_topr", ptr(G), n, r, ptr(U), ptr(w))


TDX_EIG_TRY_SUBSPACE = 1


def sym_eig_topr_ex(G, n, r, U, w=None, U0=None, subspace=True):
    """sym_eig_topr with the subspace path tried first (start block U0, FP32 n x r, or the
    built-in one); returns 1 if the subspace path produced U, 0 if the tridiagonal one did."""
    m = ctypes.c_int(-1)
    _call("tdx_sym_eig_topr_ex", ptr(G), n, r, ptr(U0), TDX_EIG_TRY_SUBSPACE if subspace else 0, ptr(U), ptr(w),
          ctypes.byref(m))
    return m.value


def ttm_ex(X, dims, mode, U, r, Y, dtype=TDX_F32):
    """Y = X x_mode U^T with U (I_mode x r) column-major; X FP32 or BF16 bits."""
    _call("tdx_ttm_ex", ptr(X), dtype, len(dims), dims_arr(dims), mode, ptr(U), r, ptr(Y))


def unfold_gram(X, dims, mode, G, dtype=TDX_F32):
    _call("tdx_unfold_gram_ex", ptr(X), dtype, len(dims), dims_arr(dims), mode, ptr(G))


def tucker_hooi(X, dims, ranks, iters, core, U, bf16_ttm=False, history=True):
    """HOSVD + `iters` HOOI iterations; returns the relative error after each iteration
    (None with history=False, which skips computing them)."""
    import numpy as np
    hist = np.zeros(max(iters, 1)) if history else None
    _call("tdx_tucker_hooi", ptr(X), len(dims), dims_arr(dims), dims_arr(ranks), iters, 1 if bf16_ttm else 0,
          ptr(core), ptr_arr(U), hist.ctypes.data if history else None)
    return hist[:iters] if history else None


class Batch:
    """Packed layout of many small 3rd-order tensors: shapes (n x 3), ranks (n), and the
    element offsets of each tensor and factor in the packed buffers."""

    def __init__(self, shapes, ranks):
        import numpy as np
        self.shapes = np.asarray(shapes, np.int64).reshape(-1, 3)
        self.ranks = np.asarray(ranks, np.int64)
        n = len(self.ranks)
        sizes = np.prod(self.shapes, axis=1)
        self.xoff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.x_floats = int(sizes.sum())
        usz = (self.shapes * self.ranks[:, None]).ravel()
        self.uoff = np.concatenate([[0], np.cumsum(usz)[:-1]]).astype(np.int64)
        self.u_floats = int(usz.sum())
        self.n = n

    def _args(self):
        return (self.shapes.ravel().ctypes.data, self.ranks.ctypes.data, self.xoff.ctypes.data,
                self.uoff.ctypes.data)


def cp_als_batched(X, U, batch, sweeps, init=0, seed=0, variant=-1, want_fit=True, want_ms=True):
    import numpy as np
    fit = np.zeros((batch.n, max(sweeps, 1)), np.float64)
    ms = np.zeros(max(sweeps, 1), np.float32)
    shapes = np.ascontiguousarray(batch.shapes.ravel())
    st = _call_status("tdx_cp_als_batched", ptr(X), ptr(U), shapes.ctypes.data, batch.ranks.ctypes.data,
                      batch.xoff.ctypes.data, batch.uoff.ctypes.data, batch.n, sweeps, init, seed, variant,
                      fit.ctypes.data if want_fit else None, ms.ctypes.data if want_ms else None)
    if st not in (TDX_OK, TDX_WARN_SINGULAR):
        raise TdxError(st, "tdx_cp_als_batched")
    return st, fit[:, :sweeps], ms[:sweeps]


def cp_batched_autotune(X, U, batch, trial_sweeps=2, cache_path=None):
    import numpy as np
    best = ctypes.c_int()
    nv = lib().tdx_cp_batched_variants()
    ms = np.zeros(nv, np.float32)
    shapes = np.ascontiguousarray(batch.shapes.ravel())
    _call("tdx_cp_batched_autotune", ptr(X), ptr(U), shapes.ctypes.data, batch.ranks.ctypes.data,
          batch.xoff.ctypes.data, batch.uoff.ctypes.data, batch.n, batch.u_floats, trial_sweeps,
          cache_path.encode() if cache_path else None, ctypes.byref(best), ms.ctypes.data)
    names = [lib().tdx_cp_batched_variant_name(v).decode() for v in range(nv)]
    return best.value, dict(zip(names, ms.tolist()))


def mlp_param_count(nin, width, nout):
    return lib().tdx_mlp_param_count(nin, width, nout)


def mlp_init(params, nin, width, nout, seed):
    _call("tdx_mlp_init", ptr(params), nin, width, nout, seed)


def mlp_forward(params, x, batch, nin, width, nout, y):
    _call("tdx_mlp_forward", ptr(params), ptr(x), batch, nin, width, nout, ptr(y))


def mlp_backward(params, x, t, batch, nin, width, nout, grad, loss=None):
    _call("tdx_mlp_backward", ptr(params), ptr(x), ptr(t), batch, nin, width, nout, ptr(grad), ptr(loss))


def mlp_adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    _call("tdx_mlp_adam", ptr(params), ptr(grad), ptr(m), ptr(v), params.numel(), step, lr, b1, b2, eps)


def mlp_train(params, X, T, nsamples, nin, width, nout, epochs, batch, lr, seed):
    import numpy as np
    hist = np.zeros(max(epochs, 1), np.float32)
    _call("tdx_mlp_train", ptr(params), ptr(X), ptr(T), nsamples, nin, width, nout, epochs, batch, lr, seed,
          hist.ctypes.data)
    return hist[:epochs]


def synchronize():
    _call("tdx_synchronize")
