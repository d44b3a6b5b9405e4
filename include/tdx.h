/*
 * tdx.h -- C ABI of the MI355X-native dense tensor-decomposition library (libtdx.so).
 *
 * Every entry point takes plain device pointers and 64-bit extents, returns an int
 * status (TDX_OK == 0), launches on the library stream (tdx_set_stream), creates no
 * per-call handles, and never frees, resets or overwrites caller memory other than the
 * declared outputs.  Storage is column-major: X[i + j*I + k*I*J (+ l*I*J*K)], factors
 * U[i + r*I] (reference convention, src/tensor2matrix.cu:4-36, src/cp_als.cu:5-6).
 *
 * Each declaration names the reference interface it replaces (inc/cudecompose.h and
 * the src/ file that implements it).  The C++ drop-in surface with the reference's
 * exact names lives in include/cudecompose.h on top of this header.
 */
#ifndef TDX_H
#define TDX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference reports none: src/mttkrp.cu:75-77 prints and returns) ---- */
enum {
    TDX_OK = 0,
    TDX_ERR_ARG = 1,          /* bad flag/mode, non-positive extent, null pointer, unsupported rank */
    TDX_ERR_ALLOC = 2,        /* device allocation failed */
    TDX_ERR_HIP = 3,          /* a HIP runtime call failed */
    TDX_WARN_SINGULAR = 4,    /* Gram-Hadamard not SPD: pseudo-inverse fallback was used */
    TDX_ERR_NONFINITE = 5,    /* NaN/Inf detected in the fit */
    TDX_ERR_UNSUPPORTED = 6,  /* shape/dtype combination not implemented */
    TDX_ERR_TOO_LARGE = 7,    /* tdx_mttkrp_ex_noalloc: panels + partials exceed its fixed scratch */
    TDX_ERR_IO = 8            /* tensor file: open/read/write failed or not a valid .tdxq file */
};

/* storage types of the tensor operand */
enum { TDX_F32 = 0, TDX_BF16 = 1, TDX_FP8_E4M3 = 2, TDX_F16 = 3 };

/* ---- library state (no reference counterpart: the reference uses the legacy default
 *      stream and per-call cuBLAS/cuSOLVER handles, e.g. src/kr.cu:7-10) ---- */
int tdx_init(void);                       /* optional: selects nothing, warms the device */
int tdx_set_stream(void *hip_stream);     /* hipStream_t; NULL = default stream */
void *tdx_get_stream(void);
int tdx_last_status(void);                /* last status recorded by any entry point */
const char *tdx_status_string(int status);
int tdx_synchronize(void);                /* hipStreamSynchronize on the library stream */
int tdx_release_workspace(void);          /* frees the internal workspace pool */
/* blocking copies on the library stream (used by printTensor, src/kr.cu:13-30) */
int tdx_copy_to_host(void *dst, const void *src, size_t bytes);
int tdx_copy_to_device(void *dst, const void *src, size_t bytes);

/* ---- elementwise and permutation primitives ---- */
/* hadmard: src/hadmard.cu:10-15 -- C = A o B over a*b floats */
int tdx_hadamard(const float *A, const float *B, float *C, long a, long b);
/* hadmard3: src/hadmard.cu:25-30 -- ABC = (A o B) o C */
int tdx_hadamard3(const float *A, const float *B, const float *C, float *ABC, long a, long b);
/* kr: src/kr.cu:3-11 -- AkrB ((a*b) x r) = A (.) B, B index fastest */
int tdx_kr(const float *A, const float *B, float *AkrB, long a, long b, long r);
/* tensor_product: src/tensorproduct.cu:4-14 -- Kronecker A (a x b) (x) B (c x d) */
int tdx_tensor_product(const float *A, const float *B, float *C, long a, long b, long c, long d);
/* t2m: src/tensor2matrix.cu:38-52 -- mode-flag unfolding copy of an m x n x k tensor */
int tdx_t2m(const float *T1, float *T2, long flag, long m, long n, long k);
/* tminust: src/tminust.cu:10-15 -- right := left - right */
int tdx_tminust(const float *left, float *right, long m);
/* initMat: src/tminust.cu:31-36 -- m x m identity */
int tdx_init_identity(float *A, long m);
/* f2h: src/float2half.cu:14-18 -- FP32 -> IEEE binary16 (RNE), bit patterns in B */
int tdx_f2h(const float *A, uint16_t *B, long num);
/* new: FP32 -> BF16 (RNE) and back */
int tdx_quantize_bf16(const float *X, uint16_t *Xq, long n);
int tdx_dequantize_bf16(const uint16_t *Xq, float *X, long n);
/* new: FP32 -> OCP FP8 E4M3FN with one FP32 scale per mode-1 slice X(i,:,...):
 * scales[i] = max|X(i,...)| / 448 (1 for an all-zero slice), Xq = e4m3(X / scales[i]).
 * n_rest = product of the other extents. */
int tdx_quantize_fp8_mode1(const float *X, uint8_t *Xq, float *scales, long I, long n_rest);
int tdx_dequantize_fp8_mode1(const uint8_t *Xq, const float *scales, float *X, long I, long n_rest);

/* ---- the quantized tensor file format (.tdxq), the step before MTTKRP: no reference
 *      counterpart beyond the in-memory conversion f2h (src/float2half.cu:14-18) ----
 * A 128-byte header (magic "TDXQ", version 1, order, storage, dims[4], n_scales, data_bytes,
 * FNV-1a 64 of the first 64 bytes), the FP8 slice scales (FP32, padded to 128 bytes), then
 * the elements column-major in the storage type: FP32, BF16 bit patterns, or E4M3FN codes
 * with one FP32 scale per mode-1 slice (the layout tdx_quantize_fp8_mode1 produces).
 * on_device = 1: X and scales are device pointers (staged through pinned buffers on the
 * library stream); 0: host pointers (no GPU needed). */
typedef struct {
    int order;
    int storage;             /* TDX_F32, TDX_BF16 or TDX_FP8_E4M3 */
    int64_t dims[4];         /* 1 past the order */
    int64_t n_scales;        /* dims[0] for FP8 storage, else 0 */
    int64_t data_bytes;
} tdx_tensor_info;
int tdx_tensor_info_read(const char *path, tdx_tensor_info *info);
int tdx_tensor_save(const char *path, const void *X, int storage, const float *scales, int order, const long *dims,
                    int on_device);
/* capacities in bytes; TDX_ERR_ARG when the file's data or scales do not fit; info may be NULL */
int tdx_tensor_load(const char *path, void *X, size_t x_capacity, float *scales, size_t scales_capacity,
                    int on_device, tdx_tensor_info *info);

/* ---- contractions ---- */
/* mttkrp: src/mttkrp.cu:3-80 -- XU = X(flag) (left (.) right) for an m x n x k tensor.
 * flag 1: left = C (k x r), right = B (n x r), XU m x r; flag 2: left = C, right = A,
 * XU n x r; flag 3: left = B, right = A, XU k x r.  FP32 in, FP32 out (the reference
 * rounds inputs to FP16; this library keeps FP32, see tdx_mttkrp_ex for BF16/FP8). */
int tdx_mttkrp(const float *X, const float *left, const float *right, float *XU,
               long flag, long m, long n, long k, long r);
/* new: fused MTTKRP of an order-2..4 tensor on mode `mode` (0-based) with the Khatri-Rao
 * rows formed on the fly.  X has storage `dtype` (TDX_F32/TDX_BF16/TDX_FP8_E4M3);
 * `scales` are the FP8 per-mode-1-slice scales (NULL otherwise).  factors[m] are
 * I_m x rank FP32 column-major (factors[mode] is not read).  out: dims[mode] x rank. */
int tdx_mttkrp_ex(const void *X, int dtype, const float *scales, int order, const long *dims,
                  const float *const *factors, long rank, int mode, float *out);
/* new, no counterpart in the reference: tdx_mttkrp_ex that allocates nothing, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, kernels and results (bit for bit) as tdx_mttkrp_ex;
 * the factor panels and split-K partials live in a fixed 64 MiB device array per GPU, held until
 * the process ends (its address is taken by tdx_init or by the first call on a device).  Returns
 * TDX_ERR_TOO_LARGE when they do not fit (e.g. a 1024^3 rank-64 MTTKRP needs 8 / 16 / 32 MiB
 * for BF16 / FP32 / FP8 storage).  Calls that overlap on one GPU (other streams or threads)
 * share that array and race; calls on the library stream are ordered. */
int tdx_mttkrp_ex_noalloc(const void *X, int dtype, const float *scales, int order, const long *dims,
                          const float *const *factors, long rank, int mode, float *out);
/* new: times `iters` launches of the fused MTTKRP kernel alone with hipEvents on the
 * library stream (panels and workspace prepared once); *avg_ms = mean launch time. */
int tdx_bench_mttkrp(const void *X, int dtype, const float *scales, int order, const long *dims,
                     const float *const *factors, long rank, int mode, int iters, float *avg_ms);
/* new, no counterpart in the reference (which runs one MTTKRP per mode, src/cp_als.cu:52,69,88):
 * the shared pass of an order-3 sweep.  M1 = X(1) (C (.) A) (I1 x rank) and M2 = X(2) (B (.) A)
 * (I2 x rank), both column-major like tdx_mttkrp_ex modes 1 and 2, from one pass over X that
 * writes T = X x_0 A^T (stored in BF16, rounded to nearest even: M2 carries that rounding) and a
 * pass over T.  BF16 or FP8-E4M3 storage (scales as for tdx_mttkrp_ex); requires I0 = 128 k
 * (BF16) or 256 k (FP8) with k in {2, 4, 8, 16}, I1 % 32 == 0 and rank <= 64; otherwise
 * TDX_ERR_UNSUPPORTED. */
int tdx_mttkrp_dtree(const void *X, int dtype, const float *scales, const long *dims,
                     const float *const *factors, long rank, float *M1, float *M2);
/* new: mean launch duration (ms, hipEvents on the library stream, `iters` launches each) of the
 * two kernels of the shared pass: the pass over X (T and the mode-1 partials) and the pass
 * over T (the mode-2 partials) */
int tdx_bench_mttkrp_dtree(const void *X, int dtype, const float *scales, const long *dims,
                           const float *const *factors, long rank, int iters, float *ms_m1, float *ms_m2);
/* ttm: src/ttm.cu:4-47 -- XU = X x_flag U with U (d x mode size) col-major;
 * outputs d x b x c, a x d x c or a x b x d.  FP32 in/out. */
int tdx_ttm(const float *X, const float *U, float *XU, long flag, long a, long b, long c, long d);
/* new: order-d TTM with a transposed factor, Y = X x_mode U^T, U (I_mode x r) */
int tdx_ttm_t(const float *X, int order, const long *dims, int mode, const float *U, long r, float *Y);
/* new: tdx_ttm_t for X stored as TDX_F32 or TDX_BF16 (for BF16, U is rounded to BF16 and the
 * products accumulate in FP32 on MFMA) */
int tdx_ttm_ex(const void *X, int dtype, int order, const long *dims, int mode, const float *U, long r, float *Y);
/* new: Gram of the mode-n unfolding, G = X(n) X(n)^T (I_n x I_n, FP64 out) */
int tdx_unfold_gram(const float *X, int order, const long *dims, int mode, double *G);
/* new: the same Gram for X stored as TDX_F32 or TDX_BF16 (FP32 accumulation on MFMA for the
 * shapes with 16-byte vectors along the contiguous index, FP64 sum of the partials) */
int tdx_unfold_gram_ex(const void *X, int dtype, int order, const long *dims, int mode, double *G);
/* new (the eigensolve of src/tucker_hosvd.cu:43-70, cusolverDnDsyevdx + copy of the last r
 * columns): U (n x r FP32, col-major) = eigenvectors of the r largest eigenvalues of the
 * symmetric FP64 n x n G, ascending, sign fixed so the largest-magnitude entry is positive;
 * w (n x 1, optional, NULL to skip) receives those r eigenvalues.  n <= 128. */
int tdx_sym_eig_topr(const double *G, long n, long r, float *U, double *w);
/* new: tdx_sym_eig_topr with a choice of method.  flags & TDX_EIG_TRY_SUBSPACE: first try
 * subspace iteration + Rayleigh-Ritz (r <= 32) from the start block U0 (n x r FP32, may be
 * NULL for a built-in one, may alias U), accepted only when certified to give the r largest
 * eigenvalues to 1e-12 relative residual; otherwise the tridiagonal solver runs.  *method
 * (host, optional) = 1 if the subspace path produced U, 0 otherwise.  Same output
 * conventions as tdx_sym_eig_topr. */
#define TDX_EIG_TRY_SUBSPACE 1
int tdx_sym_eig_topr_ex(const double *G, long n, long r, const float *U0, int flags, float *U, double *w,
                        int *method);

/* ---- reconstruction and error ---- */
/* rse: src/rse.cu:4-15 -- *error (host) = ||X - X1|| / ||X||.  Unlike the reference, X1
 * is NOT overwritten; the sums are FP64-accumulated with compensation. */
int tdx_rse(const float *X, const float *X1, long m, float *error);
/* gencp: src/rse.cu:17-30 -- rec = A (C (.) B)^T, an a x b x c tensor */
int tdx_gencp(float *rec, const float *A, const float *B, const float *C, long a, long b, long c, long r);
/* gencp4: src/rse.cu:53-73 */
int tdx_gencp4(float *rec, const float *A, const float *B, const float *C, const float *D,
               long a, long b, long c, long d, long r);
/* gentucker: src/rse.cu:32-51 -- rec = core x1 A x2 B x3 C */
int tdx_gentucker(float *rec, const float *core, const float *A, const float *B, const float *C,
                  long a, long b, long c, long r1, long r2, long r3);
/* gentucker4: src/rse.cu:75-96 */
int tdx_gentucker4(float *rec, const float *core, const float *A, const float *B, const float *C,
                   const float *D, long a, long b, long c, long d, long r1, long r2, long r3, long r4);
/* new: compensated CP residual ||X - [[U]]|| / ||X|| without forming the reconstruction */
int tdx_cp_relerr(const void *X, int dtype, const float *scales, int order, const long *dims,
                  const float *const *factors, long rank, double *relerr);

/* ---- synthetic inputs (Philox-4x32-10, identical to the oracle's stream) ---- */
int tdx_fill_uniform(float *out, long n, uint64_t seed, uint32_t stream);
int tdx_fill_normal(float *out, long n, uint64_t seed, uint32_t stream);
/* X = [[U]] + eta * ||[[U]]|| * E/||E||, E ~ N(0,1) from (seed, stream); norms in FP64 */
int tdx_gen_cp_tensor(float *X, int order, const long *dims, const float *const *factors, long rank,
                      double eta, uint64_t seed, uint32_t stream);

/* ---- CP-ALS ---- */
typedef struct {
    int max_sweeps;          /* sweeps of the main phase (reference: 500, src/cp_als.cu:42) */
    int refine_sweeps;       /* extra FP32-storage sweeps after a BF16/FP8 phase */
    int storage;             /* TDX_F32, TDX_BF16 or TDX_FP8_E4M3 for the main phase */
    int init;                /* 0: seeded U(0,1) for modes 1..d-1 (reference semantics,
                                src/cp_als.cu:20-24, Philox instead of cuRAND); 1: caller init */
    uint64_t seed;           /* init seed (reference: 1234) */
    double tol;              /* stop when |fit change| < tol (0: run all sweeps) */
    int record_fit;          /* compute the fast-fit relative error every sweep */
    int line_search;         /* extrapolation after every sweep, U_a = U_start + alpha (U - U_start),
                                accepted only if its exact relative error is lower (Tensorlab
                                cpd_als.m:234-339): 0 none (plain ALS), 1 alpha from the MLP
                                (mlp_params), 2 alpha = (sweep+1)^(1/3) (cpd_lsb.m:30),
                                3 no step, record the relative error at every grid alpha.
                                FP32 storage only: with BF16 / FP8 storage TDX_ERR_UNSUPPORTED */
    int use_graph;           /* capture each sweep in a hipGraph (default 1) */
    const float *mlp_params; /* device MLP parameters, nin = TDX_LS_FEATURES, nout = 1 */
    int mlp_width;
    int ngrid;               /* line_search == 3: number of grid points (<= 8) */
    double grid[8];
    struct tdx_cp_trace *trace; /* optional per-sweep host outputs */
    double target_relerr;    /* stop once the relative error is <= this (0: off) */
    int shared_pass;         /* order 3, BF16/FP8 storage: the mode-1 and mode-2 MTTKRPs from one
                                pass over X that also writes T = X x_0 A^T (I2 I1 Rp floats), then
                                a pass over T (default 1; no counterpart in the reference, which
                                runs one MTTKRP per mode); 0: one fused MTTKRP per mode */
    int opening_sweeps;      /* BF16/FP8 storage: the first opening_sweeps of the max_sweeps main
                                sweeps read the FP32 tensor (default 0).  The early sweeps decide
                                which components ALS resolves; rounding the Khatri-Rao operand to
                                BF16 there can settle a run on another plateau */
    int opening_split;       /* BF16 storage: the opening sweeps stream the BF16 copy with the
                                Khatri-Rao operand split into hi + lo BF16 parts (near-FP32
                                factors, BF16-rounded tensor) instead of the FP32 tensor (default
                                0; falls back to the FP32 tensor where a mode's MTTKRP is not the
                                LDS-staged kernel) */
} tdx_cp_options;

/* features of the extrapolation MLP, computed on the device after every sweep k:
 * [0] log10 relerr_k, [1..3] log10(relerr_{k-j}/relerr_{k-j+1}), [4..7] log10 of the relative
 * factor change ||U_m - U_start,m|| / ||U_m|| per mode (0 beyond the order), [8] log10(k+1),
 * [9] 1 for order 4, [10] log of the last accepted alpha (0 if none), [11] accept rate.
 * MLP output beta gives alpha = exp(clamp(beta, 0, log 32)). */
#define TDX_LS_FEATURES 12

typedef struct tdx_cp_trace {
    float *alpha;            /* [sweeps] proposed alpha (line_search 1, 2) */
    int *accepted;           /* [sweeps] 1 if the extrapolated point was kept */
    float *features;         /* [sweeps * TDX_LS_FEATURES] */
    double *grid_relerr;     /* [sweeps * ngrid] relative error at each grid alpha (line_search 3) */
    double *plain_relerr;    /* [sweeps] relative error of the plain ALS update of every sweep */
    float *elapsed_ms;       /* [sweeps] device ms from the call's first device work (norm of X,
                                conversion, panels) to the end of every sweep: time to an error
                                level including every phase change (any line_search mode) */
} tdx_cp_trace;

typedef struct {
    int sweeps_done;
    int status;              /* TDX_OK or TDX_WARN_SINGULAR */
    double relerr_fit;       /* last fast-fit relative error (FP64 identity) */
    double seconds;          /* device time of the whole call from the first device work on:
                                norm of X, the BF16 / FP8 conversion of X, factor panels and
                                every sweep (hipEvents) */
    double setup_seconds;    /* the part of `seconds` before the first sweep (norm, conversion,
                                panels); sweep_ms covers the rest */
} tdx_cp_report;

void tdx_cp_default_options(tdx_cp_options *opt);
/* new: CP-ALS of an order-3/4 FP32 tensor; factors[m] (I_m x rank) in/out.  The BF16 / FP8
 * copy of X for the main phase lives in the library's workspace pool (allocated by the first
 * call that needs it, reused after; tdx_release_workspace frees it).  rank <= 64
 * (larger ranks return TDX_ERR_UNSUPPORTED; the reference has no limit).
 * fit_hist (host, may be NULL) receives the per-sweep fast-fit relative error;
 * sweep_ms (host, may be NULL) receives the per-sweep device time. */
int tdx_cp_als_ex(const float *X, int order, const long *dims, long rank, float *const *factors,
                  const tdx_cp_options *opt, tdx_cp_report *report, double *fit_hist, float *sweep_ms);
/* cp_als: src/cp_als.cu:3-134 -- 500 FP32 sweeps, B and C drawn from the seeded generator
 * (seed 1234) as in the reference; A, B, C are outputs.  r <= 64 (as tdx_cp_als_ex). */
int tdx_cp_als(const float *X, float *A, float *B, float *C, long a, long b, long c, long r);
/* cp_als4: src/cp4_als.cu:3-186 */
int tdx_cp_als4(const float *X, float *A, float *B, float *C, float *D, long a, long b, long c, long d, long r);
/* cp_tensorcore: src/tensorcore_cp.cu:3-150 -- mixed precision: BF16 tensor storage on MFMA
 * with FP32 accumulation, FP64 solve, 2 FP32 refinement sweeps (the reference uses FP16 and
 * no refinement); caller buffers are never freed (unlike src/tensorcore_cp.cu:136-139). */
int tdx_cp_tensorcore(const float *X, float *A, float *B, float *C, long a, long b, long c, long r);

/* ---- batched CP-ALS (no reference counterpart: the reference decomposes one tensor per
 *      call, src/cp_als.cu:3) ---- */
/* CP-ALS of `ntensors` 3rd-order tensors packed in X (tensor t at element offset xoff[t],
 * column-major, extents dims[3t..3t+2] <= 128) with ranks[t] <= 16; factors packed in U (mode
 * m of tensor t at uoff[3t+m], I_m x R column-major).  Workgroups pull tensors from a
 * size-sorted queue: one persistent kernel per sweep, or (the split variants, the default)
 * one MTTKRP and one solve kernel per mode; the sweep is captured in a hipGraph.  init = 0 draws
 * U(0,1) from Philox stream (seed + t, m), 1 uses the caller's U.  variant: kernel variant
 * (see tdx_cp_batched_autotune), < 0 for the default.  fit_hist (host, ntensors x sweeps,
 * may be NULL) receives the fast-fit relative errors; sweep_ms (host, may be NULL) the device
 * time of every sweep. */
int tdx_cp_als_batched(const float *X, float *U, const long *dims, const long *ranks, const long *xoff,
                       const long *uoff, long ntensors, int sweeps, int init, uint64_t seed, int variant,
                       double *fit_hist, float *sweep_ms);
int tdx_cp_batched_variants(void);
const char *tdx_cp_batched_variant_name(int variant);
/* times trial_sweeps sweeps of every variant on a copy of U (U is not modified), returns the
 * fastest in *best and all timings in ms_per_sweep (host, may be NULL); the choice is cached
 * in cache_path ("key variant ms" lines, key = device + batch shape + variant set; NULL: no
 * cache) */
int tdx_cp_batched_autotune(const float *X, const float *U, const long *dims, const long *ranks, const long *xoff,
                            const long *uoff, long ntensors, long u_floats, int trial_sweeps, const char *cache_path,
                            int *best, float *ms_per_sweep);

/* ---- learned ALS extrapolation: the MLP (no reference counterpart; the step it
 *      replaces is the line search of Tensorlab cpd_als.m:165-166,234-339) ---- */
/* parameter count of y = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3, or -1 for a bad shape
 * (nin <= 64, width <= 128, nout <= 16); parameters are one FP32 device buffer laid out
 * [W1 (width x nin, row-major) | b1 | W2 (width x width) | b2 | W3 (nout x width) | b3] */
long tdx_mlp_param_count(long nin, long width, long nout);
int tdx_mlp_init(float *params, long nin, long width, long nout, uint64_t seed);      /* Glorot uniform, zero bias */
int tdx_mlp_forward(const float *params, const float *x, long batch, long nin, long width, long nout, float *y);
/* gradient of loss = mean((y - t)^2)/2 over batch*nout into grad (same layout as params);
 * loss (device float, may be NULL) receives the loss */
int tdx_mlp_backward(const float *params, const float *x, const float *t, long batch, long nin, long width,
                     long nout, float *grad, float *loss);
/* one Adam step (Kingma & Ba 2015), step >= 1 */
int tdx_mlp_adam(float *params, const float *grad, float *m, float *v, long n, long step, float lr, float b1,
                 float b2, float eps);
/* minibatch Adam training on device arrays X (nsamples x nin), T (nsamples x nout) */
int tdx_mlp_train(float *params, const float *X, const float *T, long nsamples, long nin, long width, long nout,
                  int epochs, long batch, float lr, uint64_t seed, float *loss_hist);

/* ---- Tucker ---- */
/* Tucker limits (all entry points below): every mode size <= 128 (the Gram kernel's block and
 * the one-workgroup FP64 eigensolve), 1 <= r_n <= mode size; larger modes return
 * TDX_ERR_UNSUPPORTED (the reference has no limit, decomposition/GPU/normal/dprice/de_tucker.cpp:31
 * runs 1120^3). */
/* tucker_hosvd: src/tucker_hosvd.cu:3-103 */
int tdx_tucker_hosvd(const float *X, float *core, float *U1, float *U2, float *U3,
                     long a, long b, long c, long r1, long r2, long r3);
/* tucker_hosvd4: src/tucker4_hosvd.cu:3-141 */
int tdx_tucker_hosvd4(const float *X, float *core, float *U1, float *U2, float *U3, float *U4,
                      long a, long b, long c, long d, long r1, long r2, long r3, long r4);
/* tucker_tensorcore: src/tucker_tensorcore.cu:3-139 -- BF16 TTM chain on MFMA, FP32 accumulation */
int tdx_tucker_tensorcore(const float *X, float *core, float *U1, float *U2, float *U3,
                          long a, long b, long c, long r1, long r2, long r3);
/* new: HOSVD init + `iters` HOOI iterations for orders 2-4
 * (decomposition/GPU/normal/tucker/tucker_hooi.cu:3-218, generalised).  bf16_ttm = 1 runs
 * the Grams of X and the first TTM of every chain from a BF16 copy of X.  relerr_hist (host,
 * may be NULL): the relative error after each iteration, from a core of the FP32 tensor
 * (an extra FP32 TTM chain per iteration when bf16_ttm = 1); the final core is always
 * computed from the FP32 tensor. */
int tdx_tucker_hooi(const float *X, int order, const long *dims, const long *ranks, int iters,
                    int bf16_ttm, float *core, float *const *U, double *relerr_hist);

#ifdef __cplusplus
}
#endif
#endif
