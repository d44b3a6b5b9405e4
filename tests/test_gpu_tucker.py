"""GPU Tucker (HOSVD / HOOI, orders 3 and 4) against the FP64 oracle.

Eigenvector signs follow one convention on both sides (largest-magnitude entry
positive), so factors compare directly; the reference's cuSOLVER signs are unspecified.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import tdx
from helpers import dev, empty, host, relnorm

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def tucker_tensor(dims, ranks, eta, seed=0):
    rng = np.random.default_rng(seed)
    core = rng.standard_normal(int(np.prod(ranks)))
    Us = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d, r in zip(dims, ranks)]
    X = O.tucker_reconstruct(dims, ranks, core, Us)
    E = rng.standard_normal(X.size)
    X = X + eta * np.linalg.norm(X) / np.linalg.norm(E) * E
    return X.astype(np.float32)


def _sym_cases():
    rng = np.random.default_rng(7)
    out = []
    for n, r in ((1, 1), (2, 1), (3, 2), (12, 5), (64, 16), (128, 32), (128, 128)):
        B = rng.standard_normal((n, n))
        out.append((f"spd{n}_{r}", B @ B.T, r, True))
    L = rng.standard_normal((128, 32))
    out.append(("lowrank", L @ L.T * 1e3 + 1e-3 * np.eye(128), 32, True))
    out.append(("indefinite", (lambda S: S + S.T)(rng.standard_normal((40, 40))), 7, True))
    # degenerate top-r sets: any orthonormal basis of the eigenspace is a valid answer
    out.append(("repeated", np.diag([1.0, 2, 2, 2, 3, 3, 5, 5]), 5, False))
    out.append(("zero", np.zeros((6, 6)), 3, False))
    return out


@pytest.mark.parametrize("sub", [False, True], ids=["tridiag", "try_subspace"])
@pytest.mark.parametrize("name,G,r,unique", _sym_cases(), ids=[c[0] for c in _sym_cases()])
def test_sym_eig_topr_against_lapack(name, G, r, unique, sub):
    """Tridiagonalisation + multisection + inverse iteration against LAPACK's eigh (the
    reference calls cusolverDnDsyevdx): eigenvalues, residuals, orthonormality, and the
    vectors themselves (same sign convention) where the top-r eigenvalues are distinct.
    With the subspace path tried first, the same checks hold whichever path answers
    (eigenvalues to 1e-11 where the subspace path does): it must take the low-rank case and
    must decline the zero matrix and the tied top-r set."""
    import torch
    n = G.shape[0]
    U = empty(n * r)
    w = torch.empty(r, dtype=torch.float64, device="cuda")
    Gd = dev(np.ascontiguousarray(G.ravel(order="F"), np.float64))
    m = 0
    if sub:
        m = tdx.sym_eig_topr_ex(Gd, n, r, U, w)
        if name in ("lowrank", "zero", "repeated"):
            assert m == (1 if name == "lowrank" else 0), name
    else:
        tdx.sym_eig_topr(Gd, n, r, U, w)
    Ug = host(U).reshape((n, r), order="F").astype(np.float64)
    wg = w.cpu().numpy()
    wl, Vl = np.linalg.eigh(G)
    wl, Vl = wl[n - r:], Vl[:, n - r:]
    scale = max(1.0, np.max(np.abs(wl)))
    assert np.max(np.abs(wg - wl)) <= (1e-11 if m else 1e-13) * scale
    assert np.max(np.abs(Ug.T @ Ug - np.eye(r))) < 1e-6
    assert np.max(np.linalg.norm(G @ Ug - Ug * wg, axis=0)) <= 1e-6 * scale
    if unique:
        for c in range(r):
            k = int(np.argmax(np.abs(Vl[:, c])))
            if Vl[k, c] < 0:
                Vl[:, c] = -Vl[:, c]
        assert np.max(np.abs(Ug - Vl)) < 1e-5


def _signed_top(G, r):
    n = G.shape[0]
    wl, Vl = np.linalg.eigh(G)
    wl, Vl = wl[n - r:], Vl[:, n - r:].copy()
    for c in range(r):
        k = int(np.argmax(np.abs(Vl[:, c])))
        if Vl[k, c] < 0:
            Vl[:, c] = -Vl[:, c]
    return wl, Vl


def _gapped_cases():
    """Grams of rank-r-plus-0.5%-noise blocks: the top r eigenvalues stand clear of the rest."""
    rng = np.random.default_rng(11)
    out = []
    for n, r, k in ((128, 32, 4096), (128, 32, 256), (64, 16, 300), (24, 5, 200), (12, 3, 50), (5, 5, 20),
                    (7, 1, 30)):
        A = rng.standard_normal((n, r)) @ rng.standard_normal((r, k))
        A += 5e-3 * np.linalg.norm(A) / np.sqrt(n * k) * rng.standard_normal((n, k))
        out.append((f"gram{n}_{r}_{k}", A @ A.T, r))
    return out


@pytest.mark.parametrize("name,G,r", _gapped_cases(), ids=[c[0] for c in _gapped_cases()])
def test_sym_eig_subspace_path(name, G, r):
    """Subspace iteration + Rayleigh-Ritz where the top-r eigenvalues stand clear of the rest:
    taken (method 1) from the built-in start, from a perturbed warm start and from a warm start
    in the output buffer itself, each matching LAPACK's eigenpairs.  Ritz values carry the
    1e-12 relative residual the path stops at, hence 1e-11 on eigenvalues (the tridiagonal
    solver reaches 1e-13)."""
    import torch
    n = G.shape[0]
    Gd = dev(np.ascontiguousarray(G.ravel(order="F"), np.float64))
    wl, Vl = _signed_top(G, r)
    scale = np.max(np.abs(wl))
    rng = np.random.default_rng(5)
    P = (Vl + 1e-3 * rng.standard_normal(Vl.shape)).ravel(order="F").astype(np.float32)
    for start in ("cold", "warm", "inplace"):
        U = dev(P) if start == "inplace" else empty(n * r)
        w = torch.empty(r, dtype=torch.float64, device="cuda")
        U0 = {"cold": None, "warm": dev(P), "inplace": U}[start]
        m = tdx.sym_eig_topr_ex(Gd, n, r, U, w, U0=U0)
        assert m == 1, (name, start)
        Ug = host(U).reshape((n, r), order="F").astype(np.float64)
        assert np.max(np.abs(w.cpu().numpy() - wl)) <= 1e-11 * scale, (name, start)
        assert np.max(np.abs(Ug - Vl)) < 1e-5, (name, start)


def test_sym_eig_subspace_declines_wrong_invariant_start():
    """A start block spanning an invariant subspace that misses the top eigenvector converges
    at once to the wrong eigenvalues; the certificate rejects them and the tridiagonal solver
    answers (method 0)."""
    import torch
    n, r = 16, 4
    G = np.diag(np.arange(n, 0, -1).astype(np.float64) ** 3)  # eigenvalues 16^3 > 15^3 > ...
    U0 = np.zeros((n, r), np.float32)
    for c in range(r):
        U0[c + 1, c] = 1.0  # e_2 .. e_5: misses e_1
    U = empty(n * r)
    w = torch.empty(r, dtype=torch.float64, device="cuda")
    m = tdx.sym_eig_topr_ex(dev(np.ascontiguousarray(G.ravel(order="F"))), n, r, U, w,
                            U0=dev(U0.ravel(order="F")))
    assert m == 0
    wl, Vl = _signed_top(G, r)
    assert np.max(np.abs(w.cpu().numpy() - wl)) <= 1e-13 * wl[-1]
    assert np.max(np.abs(host(U).reshape((n, r), order="F") - Vl)) < 1e-6


GRAM_SHAPES = [
    (64, 48, 32, 16),      # case B everywhere except mode 0 (case A, In % 8 == 0)
    (128, 40, 24),         # In = 128 in case A; P = 128 / 5120 in case B
    (32, 96, 128),         # the largest In in case B
    (20, 16, 12),          # P = 20 / 320: falls back to the LDS kernel in modes 1-2
    (18, 32, 33),          # In = 18 in case A (not a multiple of 8): fallback
]


@pytest.mark.parametrize("dims", GRAM_SHAPES, ids=["x".join(map(str, d)) for d in GRAM_SHAPES])
@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_unfold_gram_all_modes(dims, dtype):
    """G = X(n) X(n)^T against the FP64 numpy Gram of the same (dequantised) values; BF16
    products are exact in FP32, so both paths are accumulation-error only."""
    import torch
    rng = np.random.default_rng(11)
    X = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    Xd = dev(X)
    if dtype == "bf16":
        Xq = torch.empty(X.size, dtype=torch.int16, device="cuda")
        tdx.quantize_bf16(Xd, Xq)
        Xv = (host(Xq).astype(np.uint16).astype(np.uint32) << 16).view(np.float32)
        src, dt = Xq, tdx.TDX_BF16
    else:
        Xv, src, dt = X, Xd, tdx.TDX_F32
    T = Xv.astype(np.float64).reshape(dims, order="F")
    for mode in range(len(dims)):
        In = dims[mode]
        G = torch.empty(In * In, dtype=torch.float64, device="cuda")
        tdx.unfold_gram(src, dims, mode, G, dtype=dt)
        Gg = G.cpu().numpy().reshape((In, In), order="F")
        Xn = np.moveaxis(T, mode, 0).reshape(In, -1)
        Go = Xn @ Xn.T
        err = np.max(np.abs(Gg - Go)) / np.max(np.abs(Go))
        assert err < 2e-6, (dims, dtype, mode, err)
        assert np.array_equal(Gg, Gg.T)


TTM_SHAPES = [
    ((64, 48, 32, 16), 8),    # TR in modes 1-3, P == 1 in mode 0
    ((128, 40, 24), 32),
    ((32, 128, 16), 20),      # Ra not a multiple of 16
    ((16, 24, 8, 4), 40),     # Ra up to 64 (BF16 uses 4 column blocks, FP32 falls back)
    ((20, 18, 12), 5),        # P = 20 / 360: falls back in modes 1-2; In = 20 in mode 0
]


@pytest.mark.parametrize("dims,r", TTM_SHAPES, ids=["x".join(map(str, d)) + f"_r{r}" for d, r in TTM_SHAPES])
@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_ttm_transposed_all_modes(dims, r, dtype):
    """Y = X x_n U^T against FP64 numpy on the same (dequantised) values; for BF16 storage U
    is rounded to BF16 as the kernel does, so both sides multiply identical values."""
    import torch
    rng = np.random.default_rng(5)
    X = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    Xd = dev(X)
    if dtype == "bf16":
        Xq = torch.empty(X.size, dtype=torch.int16, device="cuda")
        tdx.quantize_bf16(Xd, Xq)
        Xv = (host(Xq).astype(np.uint16).astype(np.uint32) << 16).view(np.float32)
        src, dt = Xq, tdx.TDX_BF16
    else:
        Xv, src, dt = X, Xd, tdx.TDX_F32
    T = Xv.astype(np.float64).reshape(dims, order="F")
    for mode in range(len(dims)):
        In = dims[mode]
        rr = min(r, In) if dtype == "f32" else r
        U = rng.standard_normal((In, rr)).astype(np.float32)
        Ud = dev(np.ascontiguousarray(U.ravel(order="F")))
        if dtype == "bf16":
            Uq = torch.empty(U.size, dtype=torch.int16, device="cuda")
            tdx.quantize_bf16(Ud, Uq)
            Uv = (host(Uq).astype(np.uint16).astype(np.uint32) << 16).view(np.float32).reshape((In, rr), order="F")
        else:
            Uv = U
        out_dims = list(dims)
        out_dims[mode] = rr
        Y = empty(int(np.prod(out_dims)))
        tdx.ttm_ex(src, dims, mode, Ud, rr, Y, dtype=dt)
        Yo = np.moveaxis(np.tensordot(T, Uv.astype(np.float64), axes=([mode], [0])), -1, mode)
        Yg = host(Y).reshape(out_dims, order="F")
        err = np.max(np.abs(Yg - Yo)) / np.max(np.abs(Yo))
        assert err < 2e-6, (dims, rr, dtype, mode, err)


def test_gentucker_golden():
    g = GOLD["hand_tucker_3x4x2"]
    dims, ranks = g["dims"], g["ranks"]
    A, B, C, G = [dev(np.asarray(g[k], np.float32)) for k in "ABCG"]
    rec = empty(int(np.prod(dims)))
    tdx.gentucker(rec, G, A, B, C, *dims, *ranks)
    assert np.array_equal(host(rec), np.asarray(g["X"], np.float32))


@pytest.mark.parametrize("dims,ranks", [((20, 16, 12), (4, 3, 5)), ((32, 24, 16, 12), (6, 5, 4, 3))])
def test_hosvd_matches_oracle(dims, ranks):
    X = tucker_tensor(dims, ranks, 1e-3)
    Uo, co = O.tucker_hosvd(X, dims, ranks)
    Ud = [empty(d * r) for d, r in zip(dims, ranks)]
    core = empty(int(np.prod(ranks)))
    if len(dims) == 3:
        tdx.tucker_hosvd(dev(X), core, *Ud, *dims, *ranks)
    else:
        tdx.tucker_hosvd4(dev(X), core, *Ud, *dims, *ranks)
    for u, uo, d, r in zip(Ud, Uo, dims, ranks):
        ug = host(u).reshape((d, r), order="F")
        assert np.allclose(ug.T @ ug, np.eye(r), atol=1e-5)
        assert relnorm(ug, uo) < 1e-4
    assert relnorm(host(core), co) < 1e-4


HOOI_SHAPES = [((40, 30), (4, 4)), ((20, 16, 12), (4, 3, 5)), ((24, 20, 16, 12), (5, 4, 4, 3))]


@pytest.mark.parametrize("dims,ranks", HOOI_SHAPES, ids=["x".join(map(str, d)) for d, _ in HOOI_SHAPES])
def test_hooi_relerr_matches_oracle(dims, ranks):
    """HOOI (dimension-tree TTM chains, warm-started eigensolves) against the oracle's
    per-mode full chains and full eigensolves, orders 2-4: error history, core norm, factors;
    the run without history gives the same final core bit for bit."""
    X = tucker_tensor(dims, ranks, 5e-3, seed=1)
    U0, _ = O.tucker_hosvd(X, dims, ranks)
    Uo, co, hist_o = O.tucker_hooi(X, dims, ranks, 3, U0)
    Ud = [empty(d * r) for d, r in zip(dims, ranks)]
    core = empty(int(np.prod(ranks)))
    hist = tdx.tucker_hooi(dev(X), dims, ranks, 3, core, Ud)
    # relerr = sqrt(1 - ||core||^2/||X||^2) from an FP32 TTM chain: an FP32 error of ~1e-7 in
    # ||core||^2/||X||^2 moves relerr (5e-3 here) by ~1e-5
    assert np.max(np.abs(hist - hist_o)) < 2e-5
    assert abs(np.linalg.norm(host(core)) - np.linalg.norm(co)) / np.linalg.norm(co) < 1e-5
    for u, uo, d, r in zip(Ud, Uo, dims, ranks):
        assert relnorm(host(u).reshape((d, r), order="F"), uo) < 1e-4
    U2 = [empty(d * r) for d, r in zip(dims, ranks)]
    core2 = empty(int(np.prod(ranks)))
    assert tdx.tucker_hooi(dev(X), dims, ranks, 3, core2, U2, history=False) is None
    assert np.array_equal(host(core2), host(core))


def test_ttm_golden_and_random():
    g = GOLD["ttm_srand2"]
    a, b, c = g["dims"]
    d = g["d"]
    X = dev(np.asarray(g["X"], np.float32))
    for f in (1, 2, 3):
        out = empty({1: d * b * c, 2: a * d * c, 3: a * b * d}[f])
        tdx.ttm(X, dev(np.asarray(g["U"][f - 1], np.float32)), out, f, a, b, c, d)
        assert np.allclose(host(out), np.asarray(g["XU"][f - 1]), rtol=1e-6, atol=1e-7)


def test_bf16_storage_hooi_and_tensorcore_entry():
    """BF16 Grams and first TTMs (tucker_tensorcore, HOOI with bf16_ttm): the subspaces come
    from the BF16 copy, the core and the relative error from the FP32 tensor, so the error
    matches the FP64 oracle's to second order in the BF16 rounding."""
    dims, ranks = (32, 24, 16, 12), (6, 5, 4, 3)
    X = tucker_tensor(dims, ranks, 5e-3, seed=2)
    U0, _ = O.tucker_hosvd(X, dims, ranks)
    _, co, hist_o = O.tucker_hooi(X, dims, ranks, 2, U0)
    Ud = [empty(d * r) for d, r in zip(dims, ranks)]
    core = empty(int(np.prod(ranks)))
    hist = tdx.tucker_hooi(dev(X), dims, ranks, 2, core, Ud, bf16_ttm=True)
    assert np.max(np.abs(hist - hist_o)) < 2e-5
    d3, r3 = (20, 16, 12), (4, 3, 5)
    X3 = tucker_tensor(d3, r3, 1e-3, seed=3)
    Uo, co3 = O.tucker_hosvd(X3, d3, r3)
    U3 = [empty(d * r) for d, r in zip(d3, r3)]
    c3 = empty(int(np.prod(r3)))
    tdx.tucker_tensorcore(dev(X3), c3, *U3, *d3, *r3)
    for u, uo, d, r in zip(U3, Uo, d3, r3):
        assert relnorm(host(u).reshape((d, r), order="F"), uo) < 1e-2
