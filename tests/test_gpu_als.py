"""GPU CP-ALS against the FP64 oracle from identical initial factors.

Bars: the plain FP32 path tracks the FP64 oracle's per-sweep relative error within 1e-5
(C1 recipe: 128^3, rank 16, U(0,1) factors, 1% noise, U(0,1) init, 50 sweeps); BF16 and
FP8 storage followed by FP32 refinement sweeps land within 1e-4 of the oracle's relative
error; the fast-fit identity agrees with the direct compensated residual.
"""
import numpy as np
import pytest

import oracle as O
import tdx
from helpers import dev, empty, host, relnorm

pytestmark = pytest.mark.gpu


def make_cp_tensor(dims, R, seed_f=0, eta=1e-2, noise_seed=2, kind="uniform"):
    """Low-rank-plus-noise tensor built on the device by tdx_gen_cp_tensor from U(0,1)
    factors of the Philox stream (seed_f, mode), or N(0,1) factors with unit columns."""
    if kind == "uniform":
        F = [O.uniform(d * R, seed_f, m).reshape((d, R), order="F") for m, d in enumerate(dims)]
    else:
        F = [O.normal(d * R, seed_f, m).reshape((d, R), order="F") for m, d in enumerate(dims)]
        F = [(f / np.linalg.norm(f.astype(np.float64), axis=0)).astype(np.float32) for f in F]
    X = empty(int(np.prod(dims)))
    tdx.gen_cp_tensor(X, dims, [dev(f.ravel(order="F")) for f in F], R, eta, noise_seed, 0)
    return X, F


def init_factors(dims, R, seed=1):
    return [O.uniform(d * R, seed, m).reshape((d, R), order="F") for m, d in enumerate(dims)]


def run_gpu_als(X, dims, R, U0, sweeps, storage=tdx.TDX_F32, refine=0, graph=True):
    Ud = [dev(u.ravel(order="F")) for u in U0]
    o = tdx.default_cp_options()
    o.max_sweeps = sweeps
    o.refine_sweeps = refine
    o.storage = storage
    o.init = 1
    o.record_fit = 1
    o.use_graph = 1 if graph else 0
    st, rep, fit, ms = tdx.cp_als_ex(X, dims, R, Ud, o)
    return st, rep, fit, [host(u).reshape((d, R), order="F") for u, d in zip(Ud, dims)]


def test_gen_cp_tensor_matches_recipe():
    dims, R = (24, 16, 20), 3
    X, F = make_cp_tensor(dims, R, eta=1e-2)
    Xh = host(X).astype(np.float64)
    X0 = O.cp_reconstruct(dims, [f.astype(np.float64) for f in F])
    E = Xh - X0
    assert np.isclose(np.linalg.norm(E) / np.linalg.norm(X0), 1e-2, rtol=1e-4)
    zo = O.normal(Xh.size, 2, 0).astype(np.float64)
    assert abs(np.dot(E, zo) / (np.linalg.norm(E) * np.linalg.norm(zo)) - 1) < 1e-6


def test_c1_plain_fp32_tracks_fp64_oracle():
    dims, R, sweeps = (128, 128, 128), 16, 50
    X, _ = make_cp_tensor(dims, R, eta=1e-2)
    U0 = init_factors(dims, R, seed=1)
    st, rep, fit, U = run_gpu_als(X, dims, R, U0, sweeps)
    assert st == tdx.TDX_OK and rep.sweeps_done == sweeps
    Xh = host(X)
    Uo, fit_o, dr_o, st_o = O.cp_als(Xh, dims, U0, sweeps, direct=True)
    # per-sweep fit identity of the GPU vs the oracle's direct residual
    assert np.max(np.abs(fit - dr_o)) <= 1e-5, np.max(np.abs(fit - dr_o))
    # final factors: the GPU direct residual of its own factors vs the oracle
    rel_gpu = O.cp_relerr(Xh, dims, [u.astype(np.float64) for u in U])
    assert abs(rel_gpu - dr_o[-1]) <= 1e-5
    for u, uo in zip(U, Uo):
        assert relnorm(u, uo) < 1e-3
    # ALS error never increases (U(0,1) factors: a slow "swamp", far from the 1% floor at 50 sweeps)
    assert np.all(np.diff(fit) <= 1e-6)


@pytest.mark.parametrize("R", [24, 40, 64])
def test_larger_ranks_track_fp64_oracle(R):
    """Ranks past 16 and 32 take the 8- and 16-slot variants of the FP64 inverse."""
    dims, sweeps = (96, 80, 72), 12
    X, _ = make_cp_tensor(dims, R, eta=1e-2, kind="normal")
    U0 = init_factors(dims, R, seed=7)
    st, rep, fit, U = run_gpu_als(X, dims, R, U0, sweeps)
    assert st == tdx.TDX_OK and rep.sweeps_done == sweeps
    _, _, dr_o, _ = O.cp_als(host(X), dims, U0, sweeps, direct=True)
    assert np.max(np.abs(fit - dr_o)) <= 1e-4, np.max(np.abs(fit - dr_o))


def test_fit_identity_vs_direct_residual_on_device():
    dims, R = (64, 48, 40), 8
    X, _ = make_cp_tensor(dims, R, eta=1e-3)
    U0 = init_factors(dims, R, seed=3)
    st, rep, fit, U = run_gpu_als(X, dims, R, U0, 20)
    direct = tdx.cp_relerr(X, dims, [dev(u.ravel(order="F")) for u in U], R)
    assert abs(direct - fit[-1]) < 1e-6
    assert abs(direct - O.cp_relerr(host(X), dims, [u.astype(np.float64) for u in U])) < 1e-7


@pytest.mark.parametrize("storage", [tdx.TDX_F32, tdx.TDX_BF16])
def test_fit_recording_does_not_change_the_iterate(storage):
    """With the fit off, the last mode's Gram comes from k_gram_psum instead of the kernel
    that also appends the fit: the factors must be bit-identical either way."""
    dims, R = (256, 64, 48), 16
    X, _ = make_cp_tensor(dims, R, eta=1e-2, kind="normal")
    U0 = init_factors(dims, R, seed=11)
    out = []
    for rec in (1, 0):
        Ud = [dev(u.ravel(order="F")) for u in U0]
        o = tdx.default_cp_options()
        o.max_sweeps, o.refine_sweeps, o.storage, o.init, o.record_fit = 6, 1, storage, 1, rec
        st, rep, fit, ms = tdx.cp_als_ex(X, dims, R, Ud, o)
        assert st == tdx.TDX_OK
        out.append([host(u) for u in Ud])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dims,R,storage", [((32, 40, 24), 6, tdx.TDX_F32), ((32, 40, 24), 24, tdx.TDX_F32),
                                            ((256, 64, 40), 24, tdx.TDX_BF16)])
def test_graph_and_eager_identical(dims, R, storage):
    """Replayed graph and eager launches give identical fits and factors: rank <= 16 with the
    inverse in line, rank 24 with the inverse on the side stream (FP32 per-mode MTTKRP, and
    BF16 storage with the shared pass)."""
    X, _ = make_cp_tensor(dims, R)
    U0 = init_factors(dims, R, seed=4)
    _, _, f1, U1 = run_gpu_als(X, dims, R, U0, 8, storage=storage, graph=True)
    _, _, f2, U2 = run_gpu_als(X, dims, R, U0, 8, storage=storage, graph=False)
    assert np.array_equal(f1, f2)
    for a, b in zip(U1, U2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("storage", [tdx.TDX_BF16, tdx.TDX_FP8_E4M3])
def test_mixed_precision_reaches_oracle(storage):
    # converging problem (N(0,1) unit-column factors, as in the 1024^3 workload)
    dims, R = (128, 96, 64), 16
    X, _ = make_cp_tensor(dims, R, eta=1e-3, kind="normal")
    U0 = init_factors(dims, R, seed=5)
    st, rep, fit, U = run_gpu_als(X, dims, R, U0, 25, storage=storage, refine=2)
    Xh = host(X)
    _, _, dr_o, _ = O.cp_als(Xh, dims, U0, 27, direct=True)
    rel = O.cp_relerr(Xh, dims, [u.astype(np.float64) for u in U])
    assert abs(rel - dr_o[-1]) <= 1e-4, (rel, dr_o[-1])


@pytest.mark.parametrize("storage,dims", [(tdx.TDX_BF16, (256, 96, 64)), (tdx.TDX_FP8_E4M3, (512, 64, 48))])
@pytest.mark.parametrize("shared", [1, 0])
def test_mixed_sweeps_reach_oracle_with_and_without_shared_pass(storage, dims, shared):
    """Order 3, BF16/FP8 storage, modes 1 and 2 from the shared pass (T = X x_0 A^T, A kept
    BF16-exact) or from one fused MTTKRP per mode: 25 sweeps + 2 FP32 refinement sweeps end
    at the FP64 oracle's relative error to 1e-4."""
    R = 16
    X, _ = make_cp_tensor(dims, R, eta=1e-3, kind="normal")
    U0 = init_factors(dims, R, seed=5)
    Ud = [dev(u.ravel(order="F")) for u in U0]
    o = tdx.default_cp_options()
    o.max_sweeps, o.refine_sweeps, o.storage, o.init, o.record_fit = 25, 2, storage, 1, 1
    o.shared_pass = shared
    st, rep, fit, ms = tdx.cp_als_ex(X, dims, R, Ud, o)
    assert st == tdx.TDX_OK and rep.sweeps_done == 27
    U = [host(u).reshape((d, R), order="F") for u, d in zip(Ud, dims)]
    Xh = host(X)
    _, _, dr_o, _ = O.cp_als(Xh, dims, U0, 27, direct=True)
    rel = O.cp_relerr(Xh, dims, [u.astype(np.float64) for u in U])
    assert abs(rel - dr_o[-1]) <= 1e-4, (rel, dr_o[-1])


def test_shared_pass_fast_fit_is_the_iterates_error():
    """BF16 storage with the shared pass: A is BF16-exact, so the normal equations, the Grams
    and the pass over X all describe the same iterate, and the per-sweep fast-fit identity
    (FP64) matches the direct relative error of the returned factors against the FP32 tensor
    (FP64 oracle) to 5e-5 plus the effect of storing T in BF16 on <M2, C>, at a transient and
    on the plateau.  Rounding to nearest leaves each T entry off by a fraction of itself within
    +-2^-9, close to uniform and independent over entries, so the fit moves by about
    2^-9 / sqrt(3) sqrt(sum (T B C)^2) / (||X||^2 relerr) (RMS; T = X x_0 A^T of the returned
    A); the test allows six times that."""
    dims, R = (512, 64, 48), 16
    X, _ = make_cp_tensor(dims, R, eta=1e-3, kind="normal")
    U0 = init_factors(dims, R, seed=5)
    Xh = host(X)
    for k in (4, 25):
        Ud = [dev(u.ravel(order="F")) for u in U0]
        o = tdx.default_cp_options()
        o.max_sweeps, o.refine_sweeps, o.storage, o.init, o.record_fit = k, 0, tdx.TDX_BF16, 1, 1
        st, rep, fit, ms = tdx.cp_als_ex(X, dims, R, Ud, o)
        U = [host(u).reshape((d, R), order="F").astype(np.float64) for u, d in zip(Ud, dims)]
        true = O.cp_relerr(Xh, dims, U)
        X3 = Xh.reshape(dims, order="F").astype(np.float64)
        T = np.einsum("abc,ar->bcr", X3, U[0])
        rms = 2.0 ** -9 / np.sqrt(3.0) * np.sqrt(np.einsum("bcr,br,cr->", T * T, U[1] * U[1], U[2] * U[2]))
        bound = 6.0 * rms / (np.sum(X3 * X3) * true)
        assert bound <= 1e-3, bound
        assert abs(fit[-1] - true) <= 5e-5 + bound, (k, fit[-1], true, bound)
        # the factor A holds BF16 values only
        a = U[0].astype(np.float32).ravel()
        assert np.array_equal(a, (O.bf16_bits(a).astype(np.uint32) << 16).view(np.float32))


@pytest.mark.parametrize("R", [24, 40, 56])
def test_shared_pass_mode2_inverse_and_gram(R):
    """Ranks 17..64 (padded to 32, 48, 64) with the shared pass: mode 1's Gram is summed and mode 2's inverse formed
    inside the pass over T.  The per-sweep fast fit (which uses that Gram) equals the direct
    relative error of the returned factors (FP64 oracle) to 5e-5, and the sweeps follow the ones
    without the shared pass (a fused MTTKRP per mode, separate inverse) to within BF16 rounding."""
    dims = (256, 96, 64)
    X, _ = make_cp_tensor(dims, R, eta=1e-3, kind="normal")
    U0 = init_factors(dims, R, seed=7)
    Xh = host(X)
    res = {}
    for shared in (1, 0):
        Ud = [dev(u.ravel(order="F")) for u in U0]
        o = tdx.default_cp_options()
        o.max_sweeps, o.refine_sweeps, o.storage, o.init, o.record_fit = 4, 0, tdx.TDX_BF16, 1, 1
        o.shared_pass = shared
        st, rep, fit, ms = tdx.cp_als_ex(X, dims, R, Ud, o)
        assert st == tdx.TDX_OK and rep.sweeps_done == 4
        U = [host(u).reshape((d, R), order="F").astype(np.float64) for u, d in zip(Ud, dims)]
        res[shared] = (np.asarray(fit), O.cp_relerr(Xh, dims, U))
    fit1, true1 = res[1]
    assert np.all(np.isfinite(fit1))
    assert abs(fit1[-1] - true1) <= 5e-5, (fit1[-1], true1)
    fit0, true0 = res[0]
    assert not np.array_equal(fit1, fit0)  # the shared pass ran (another summation order)
    assert np.max(np.abs(fit1 - fit0)) <= 1e-2, (fit1, fit0)
    assert abs(true1 - true0) <= 1e-2, (true1, true0)


def test_order4_cp_als():
    dims, R = (24, 16, 20, 8), 5
    X, _ = make_cp_tensor(dims, R, eta=1e-3)
    U0 = init_factors(dims, R, seed=6)
    st, rep, fit, U = run_gpu_als(X, dims, R, U0, 30)
    _, _, dr_o, _ = O.cp_als(host(X), dims, U0, 30, direct=True)
    assert np.max(np.abs(fit - dr_o)) <= 1e-5


def test_exact_factors_fixed_point():
    dims, R = (40, 32, 24), 4
    X, F = make_cp_tensor(dims, R, eta=0.0)
    st, rep, fit, U = run_gpu_als(X, dims, R, F, 3)
    # the fast-fit identity cancels at this level (FP32 MTTKRP); the direct residual does not
    direct = tdx.cp_relerr(X, dims, [dev(u.ravel(order="F")) for u in U], R)
    assert direct < 2e-6 and fit[-1] < 2e-4


def test_singular_normal_equations_fall_back():
    """Duplicate factor columns make the Gram-Hadamard singular: the Cholesky fails, the
    Jacobi pseudo-inverse is used and the call reports TDX_WARN_SINGULAR."""
    dims, R = (32, 24, 16), 4
    X, _ = make_cp_tensor(dims, R)
    U0 = init_factors(dims, R, seed=7)
    for u in U0:
        u[:, 1] = u[:, 0]
    st, rep, fit, U = run_gpu_als(X, dims, R, U0, 2)
    assert st == tdx.TDX_WARN_SINGULAR
    assert np.all(np.isfinite(fit))


def test_reference_entry_points():
    dims, R = (32, 24, 16), 4
    X, F = make_cp_tensor(dims, R, eta=0.0)
    A, B, C = [empty(d * R) for d in dims]
    st = tdx.cp_als(X, A, B, C, *dims, R)
    assert st in (tdx.TDX_OK, tdx.TDX_WARN_SINGULAR)
    rec = empty(int(np.prod(dims)))
    tdx.gencp(rec, A, B, C, *dims, R)
    assert tdx.rse(X, rec, rec.numel()) < 1e-3
    # gencp against the oracle reconstruction
    Fd = [dev(f.ravel(order="F")) for f in F]
    tdx.gencp(rec, *Fd, *dims, R)
    ref = O.cp_reconstruct(dims, [f.astype(np.float64) for f in F])
    assert relnorm(host(rec), ref) < 1e-6


def test_nonfinite_tensor_reports_status():
    """A NaN in X (and an Inf) makes the fit non-finite: the driver returns TDX_ERR_NONFINITE
    instead of a silently wrong factorisation (the reference ignores every status,
    src/mttkrp.cu:75-77)."""
    dims, R = (32, 24, 16), 4
    for bad in (float("nan"), float("inf")):
        X, _ = make_cp_tensor(dims, R)
        Xh = host(X)
        Xh[1234] = bad
        U0 = init_factors(dims, R, seed=9)
        Ud = [dev(u.ravel(order="F")) for u in U0]
        o = tdx.default_cp_options()
        o.max_sweeps, o.init, o.record_fit = 3, 1, 1
        with pytest.raises(tdx.TdxError) as e:
            tdx.cp_als_ex(dev(Xh), dims, R, Ud, o)
        assert e.value.status == tdx.TDX_ERR_NONFINITE
        assert tdx.lib().tdx_last_status() == tdx.TDX_ERR_NONFINITE


def test_opening_split_is_als_of_the_bf16_tensor_with_fp32_factors():
    """BF16 storage with the split (hi + lo BF16) Khatri-Rao block, the schedule's cheaper
    opening sweeps: it is the ALS of the BF16-rounded tensor with near-FP32 factors, so its
    factors after 6 sweeps match the plain FP32 path run on the dequantised tensor (relative
    5e-5: the block is exact to 2^-17) and their direct errors against that tensor agree to 2e-7 -- where the plain BF16
    sweeps (factor operand rounded to 2^-9) miss both by orders of magnitude more.  Orders 3
    and 4, rank 16 and 8 (the transposed mode-0 kernel and the n > 0 kernel)."""
    import torch
    for dims, R in (((256, 96, 80), 16), ((64, 48, 40, 24), 8)):
        X, _ = make_cp_tensor(dims, R, eta=1e-2)
        Xb = torch.empty(X.numel(), dtype=torch.int16, device="cuda")
        tdx.quantize_bf16(X, Xb)
        Xd = empty(X.numel())
        tdx.lib().tdx_dequantize_bf16(tdx.ptr(Xb), tdx.ptr(Xd), X.numel())
        U0 = init_factors(dims, R, seed=3)
        runs = {}
        for name, (Xs, storage, opening, split) in {"split": (X, tdx.TDX_BF16, 6, 1),
                                                     "ref": (Xd, tdx.TDX_F32, 0, 0),
                                                     "bf16": (X, tdx.TDX_BF16, 0, 0)}.items():
            Ud = [dev(u.ravel(order="F")) for u in U0]
            o = tdx.default_cp_options()
            o.max_sweeps, o.refine_sweeps, o.storage, o.init = 6, 0, storage, 1
            o.opening_sweeps, o.opening_split, o.shared_pass = opening, split, 0
            tdx.cp_als_ex(Xs, dims, R, Ud, o)
            runs[name] = (Ud, tdx.cp_relerr(Xd, dims, Ud, R))
        ds = max(relnorm(host(a), host(b)) for a, b in zip(runs["split"][0], runs["ref"][0]))
        db = max(relnorm(host(a), host(b)) for a, b in zip(runs["bf16"][0], runs["ref"][0]))
        es, eb = abs(runs["split"][1] - runs["ref"][1]), abs(runs["bf16"][1] - runs["ref"][1])
        print(dims, R, "factor diff split", ds, "bf16", db, "relerr diff split", es, "bf16", eb)
        assert ds <= 5e-5 and es <= 2e-7
        assert db > 10 * ds


def test_fp8_storage_split_opening_is_the_bf16_split_opening():
    """FP8 storage with split opening sweeps streams a BF16 copy made in the same pass as
    ||X||^2 and the FP8 slice maxima; with every sweep an opening sweep its factors are those
    of the BF16-storage split opening, bit for bit.  With 2 opening sweeps and 4 FP8 sweeps
    after them the run stays within the FP8 rounding of the FP32 path's error."""
    for dims, R in (((256, 96, 80), 16), ((128, 64, 40, 24), 8)):
        X, _ = make_cp_tensor(dims, R, eta=1e-2)
        U0 = init_factors(dims, R, seed=5)
        res = {}
        for name, (storage, opening) in {"fp8": (tdx.TDX_FP8_E4M3, 6), "bf16": (tdx.TDX_BF16, 6),
                                         "fp8_main": (tdx.TDX_FP8_E4M3, 2), "fp32": (tdx.TDX_F32, 0)}.items():
            Ud = [dev(u.ravel(order="F")) for u in U0]
            o = tdx.default_cp_options()
            o.max_sweeps, o.refine_sweeps, o.storage, o.init = 6, 0, storage, 1
            o.opening_sweeps, o.opening_split = opening, 1
            tdx.cp_als_ex(X, dims, R, Ud, o)
            res[name] = ([host(u) for u in Ud], tdx.cp_relerr(X, dims, Ud, R))
        for a, b in zip(res["fp8"][0], res["bf16"][0]):
            assert np.array_equal(a, b)
        assert np.isfinite(res["fp8_main"][1])
        assert abs(res["fp8_main"][1] - res["fp32"][1]) < 2e-2, (res["fp8_main"][1], res["fp32"][1])


def test_plan_reuse_across_shapes_is_bit_identical():
    """The device buffers of a CP-ALS call are kept for later calls whose shape fits them: a
    smaller call after a larger one (other extents and rank, orders 3 and 4) gives the factors and
    fits of the same call on fresh buffers, bit for bit."""
    cases = [((40, 36, 32), 8), ((24, 20, 16), 4), ((20, 16, 12, 10), 5), ((12, 10, 8, 6), 3)]
    data = []
    for dims, R in cases:
        X, _ = make_cp_tensor(dims, R)
        data.append((X, dims, R, init_factors(dims, R, seed=11)))

    def run(X, dims, R, U0):
        Ud = [dev(u.ravel(order="F")) for u in U0]
        o = tdx.default_cp_options()
        o.max_sweeps, o.init, o.record_fit = 5, 1, 1
        st, rep, fit, _ = tdx.cp_als_ex(X, dims, R, Ud, o)
        return fit, [host(u) for u in Ud]

    fresh = []
    for case in data:
        tdx.lib().tdx_release_workspace()
        fresh.append(run(*case))
    tdx.lib().tdx_release_workspace()
    for (X, dims, R, U0), (f0, u0) in zip(data, fresh):  # large, then small: the small ones reuse
        f1, u1 = run(X, dims, R, U0)
        assert np.array_equal(f1, f0)
        assert all(np.array_equal(a, b) for a, b in zip(u1, u0))
