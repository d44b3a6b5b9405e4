"""Pins the FP64 CPU oracle before it is trusted as the checker.

Each oracle function is checked against something it does not share code with:
the golden fixtures (exact integer/Fraction arithmetic on the reference tests' own
inputs), published known-answer vectors (Philox), numpy/LAPACK closed forms, textbook
invariants and brute force on tiny inputs.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def F(a):
    return np.asarray(a, np.float32)


# ---------------------------------------------------------------- generator
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_and_normal_streams():
    u = O.uniform(200000, 7, 3)
    assert u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 5e-3
    # 23-bit grid centred in its bucket
    m = u.astype(np.float64) * 2 ** 23 - 0.5
    assert np.array_equal(m, np.round(m))
    z = O.normal(200000, 7, 3)
    assert abs(z.mean()) < 1e-2 and abs(z.std() - 1) < 1e-2
    # definition: Box-Muller of the uniform pair (2n, 2n+1)
    uu = O.uniform(20, 7, 3).astype(np.float64)
    ref = np.sqrt(-2 * np.log(uu[0::2])) * np.cos(2 * np.pi * uu[1::2])
    assert np.allclose(O.normal(10, 7, 3), ref.astype(np.float32), rtol=1e-6, atol=1e-6)


# ---------------------------------------------------------------- permutations
def test_t2m_golden_and_numpy():
    g = GOLD["t2m_srand2"]
    m, n, k = g["dims"]
    T = F(g["T"])
    for f in (1, 2, 3):
        assert np.array_equal(O.t2m(T, f, m, n, k), F(g["unfold"][f - 1]))
    rng = np.random.default_rng(0)
    for (m, n, k) in [(5, 7, 3), (1, 4, 9), (16, 1, 2)]:
        X = rng.standard_normal((m, n, k)).astype(np.float32)
        for f in (1, 2, 3):
            ref = np.moveaxis(X, f - 1, 0).reshape(X.shape[f - 1], -1, order="F").ravel(order="F")
            assert np.array_equal(O.t2m(X.ravel(order="F"), f, m, n, k), ref)


def test_kr_hadamard_kron_exact():
    rng = np.random.default_rng(1)
    a, b, r = 5, 4, 3
    A = rng.standard_normal((a, r)).astype(np.float32)
    B = rng.standard_normal((b, r)).astype(np.float32)
    out = O.kr(A.ravel(order="F"), B.ravel(order="F"), a, b, r).reshape((a * b, r), order="F")
    for t in range(r):
        assert np.array_equal(out[:, t], np.kron(A[:, t], B[:, t]))
    x, y, z = [rng.standard_normal(37).astype(np.float32) for _ in range(3)]
    assert np.array_equal(O.hadamard(x, y), x * y)
    assert np.array_equal(O.hadamard3(x, y, z), (x * y) * z)
    assert np.array_equal(O.tminust(x, y), x - y)
    A2 = rng.standard_normal((2, 3)).astype(np.float32)
    B2 = rng.standard_normal((4, 5)).astype(np.float32)
    got = O.tensor_product(A2.ravel(order="F"), B2.ravel(order="F"), 2, 3, 4, 5).reshape((8, 15), order="F")
    assert np.array_equal(got, np.kron(A2, B2))
    assert np.array_equal(O.identity(5).reshape(5, 5), np.eye(5, dtype=np.float32))


# ---------------------------------------------------------------- formats
def _bf16_numpy(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7fffffff) > 0x7f800000
    r = ((u + 0x7fff + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def test_bf16_rounding():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.standard_normal(2000).astype(np.float32),
                        F([0.0, -0.0, 1.0, np.inf, -np.inf, 3.4e38, 1e-40, -1e-45]),
                        # exact ties: mantissa low half 0x8000 with even and odd upper halves
                        np.array([0x3f808000, 0x3f818000, 0xbf808000], np.uint32).view(np.float32)])
    assert np.array_equal(O.bf16_bits(x), _bf16_numpy(x))
    ties = np.array([0x3f808000, 0x3f818000], np.uint32).view(np.float32)
    assert list(O.bf16_bits(ties)) == [0x3f80, 0x3f82]


def test_f16_matches_ieee():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(3000).astype(np.float32) * 100,
                        rng.standard_normal(1000).astype(np.float32) * 1e-6,
                        F([65504, 65519, 65520, 1e9, -1e9, 2 ** -24, 2 ** -25, 3 * 2 ** -26, 0.0])])
    assert np.array_equal(O.f16_bits(x), x.astype(np.float16).view(np.uint16))


def _e4m3_value(q):
    s = -1.0 if q & 0x80 else 1.0
    e, m = (q >> 3) & 15, q & 7
    if e == 15 and m == 7:
        return np.nan
    if e == 0:
        return s * m * 2.0 ** -9
    return s * (1 + m / 8) * 2.0 ** (e - 7)


def test_e4m3_decode_encode():
    vals = np.array([_e4m3_value(q) for q in range(256)])
    dec = O.e4m3_decode(np.arange(256, dtype=np.uint8))
    assert np.array_equal(np.isnan(dec), np.isnan(vals))
    ok = ~np.isnan(vals)
    assert np.array_equal(dec[ok], vals[ok].astype(np.float32))
    assert np.nanmax(vals) == 448.0
    # every finite code round-trips (both zeros map to their own sign)
    codes = np.arange(256, dtype=np.uint8)[ok]
    assert np.array_equal(O.e4m3_bits(dec[ok]), codes)
    # nearest-even at midpoints, saturation
    pos = np.sort(np.unique(vals[ok & (vals >= 0)]))
    mids = ((pos[:-1] + pos[1:]) / 2).astype(np.float32)
    got = O.e4m3_decode(O.e4m3_bits(mids))
    for mid, lo, hi, g in zip(mids, pos[:-1], pos[1:], got):
        lo_code = O.e4m3_bits(F([lo]))[0]
        expect = lo if lo_code % 2 == 0 else hi
        assert g == np.float32(expect)
    assert O.e4m3_decode(O.e4m3_bits(F([500.0, -1e6, 463.9])))[0:3].tolist() == [448.0, -448.0, 448.0]


def test_fp8_slice_scales():
    rng = np.random.default_rng(4)
    I, rest = 6, 50
    X = rng.standard_normal(I * rest).astype(np.float32)
    X[2::I] = 0.0  # an all-zero slice
    Q, s = O.quantize_e4m3_mode1(X, I)
    Xm = X.reshape((I, rest), order="F")
    amax = np.abs(Xm).max(axis=1)
    assert np.array_equal(s, np.where(amax > 0, amax / np.float32(448), 1).astype(np.float32))
    D = O.e4m3_decode(Q).reshape((I, rest), order="F") * s[:, None]
    rel = np.abs(D - Xm) / np.maximum(np.abs(Xm), s[:, None] * 2 ** -9)
    assert rel.max() <= 2 ** -4 + 1e-6


# ---------------------------------------------------------------- contractions
def test_mttkrp_golden_exact():
    for key in ("mttkrp_srand2", "hand_cp_3x4x2"):
        g = GOLD[key]
        dims, r = g["dims"], g["rank"]
        U = [np.asarray(g[n], np.float64).reshape((d, r), order="F") for n, d in zip("ABC", dims)]
        for mode in range(3):
            got = O.mttkrp(F(g["X"]), dims, U, mode)
            assert np.array_equal(got.ravel(order="F"), np.asarray(g["M"][mode])), (key, mode)
    g = GOLD["hand_cp_3x2x2"]
    U = [np.zeros((3, 2)), np.asarray(g["B"], float).reshape((2, 2), order="F"),
         np.asarray(g["C"], float).reshape((2, 2), order="F")]
    assert np.array_equal(O.mttkrp(F(g["X"]), g["dims"], U, 0).ravel(order="F"), np.asarray(g["M0"]))


def test_mttkrp_einsum_order3_and_4():
    rng = np.random.default_rng(5)
    dims = (5, 6, 7)
    R = 3
    X = rng.standard_normal(dims).astype(np.float32)
    U = [rng.standard_normal((d, R)) for d in dims]
    subs = ["ijk,jr,kr->ir", "ijk,ir,kr->jr", "ijk,ir,jr->kr"]
    for mode in range(3):
        ref = np.einsum(subs[mode], X.astype(np.float64), *[U[m] for m in range(3) if m != mode])
        assert np.allclose(O.mttkrp(X.ravel(order="F"), dims, U, mode), ref, rtol=1e-12, atol=1e-12)
    dims = (3, 4, 5, 2)
    X = rng.standard_normal(dims).astype(np.float32)
    U = [rng.standard_normal((d, R)) for d in dims]
    subs = ["ijkl,jr,kr,lr->ir", "ijkl,ir,kr,lr->jr", "ijkl,ir,jr,lr->kr", "ijkl,ir,jr,kr->lr"]
    for mode in range(4):
        ref = np.einsum(subs[mode], X.astype(np.float64), *[U[m] for m in range(4) if m != mode])
        assert np.allclose(O.mttkrp(X.ravel(order="F"), dims, U, mode), ref, rtol=1e-12, atol=1e-12)


def test_mttkrp_scaled_is_dequantised():
    rng = np.random.default_rng(6)
    dims = (4, 5, 6)
    X = rng.standard_normal(dims).astype(np.float32).ravel(order="F")
    Q, s = O.quantize_e4m3_mode1(X, 4)
    D = (O.e4m3_decode(Q).reshape((4, -1), order="F") * s[:, None]).ravel(order="F")
    U = [rng.standard_normal((d, 3)) for d in dims]
    for mode in range(3):
        a = O.mttkrp(D.astype(np.float32), dims, U, mode)
        b = O.mttkrp(O.e4m3_decode(Q), dims, U, mode, scales=s)
        assert np.allclose(a, b, rtol=1e-6, atol=1e-6)


def test_ttm_golden():
    g = GOLD["ttm_srand2"]
    a, b, c = g["dims"]
    d = g["d"]
    X = np.asarray(g["X"], np.float64)
    for f in (1, 2, 3):
        got = O.ttm(X, np.asarray(g["U"][f - 1], np.float64), f, a, b, c, d)
        assert np.allclose(got, np.asarray(g["XU"][f - 1]), rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------- dense linear algebra
def test_sym_eig_and_solves():
    rng = np.random.default_rng(7)
    A = rng.standard_normal((20, 20))
    S = A @ A.T + 0.1 * np.eye(20)
    w, V = O.sym_eig(S)
    assert np.allclose(w, np.linalg.eigvalsh(S), rtol=1e-12, atol=1e-10)
    assert np.allclose(V.T @ V, np.eye(20), atol=1e-12)
    assert np.allclose(V @ np.diag(w) @ V.T, S, atol=1e-10)
    for c in range(20):  # sign convention
        assert V[np.argmax(np.abs(V[:, c])), c] > 0
    M = rng.standard_normal((30, 20))
    st, U = O.cholesky_solve_rows(S, M)
    assert st == 0
    assert np.allclose(U, np.linalg.solve(S, M.T).T, rtol=1e-10, atol=1e-10)
    # singular V: duplicate columns -> pseudo-inverse branch, minimum-norm solution
    B = rng.standard_normal((20, 5))
    B = np.hstack([B, B[:, :1]])
    Vs = B.T @ B
    Ms = rng.standard_normal((4, 6)) @ Vs
    st, U = O.cholesky_solve_rows(Vs, Ms)
    assert st != 0
    assert np.allclose(U, Ms @ np.linalg.pinv(Vs, rcond=1e-12), atol=1e-8)


# ---------------------------------------------------------------- CP
def test_cp_invariants():
    rng = np.random.default_rng(8)
    dims = (9, 10, 11)
    R = 3
    Ftrue = [rng.random((d, R)) for d in dims]
    X = O.cp_reconstruct(dims, Ftrue).astype(np.float32)
    # gencp golden
    g = GOLD["hand_cp_3x4x2"]
    Ug = [np.asarray(g[n], float).reshape((d, 2), order="F") for n, d in zip("ABC", g["dims"])]
    assert np.array_equal(O.cp_reconstruct(g["dims"], Ug), np.asarray(g["gencp"]))
    # exact factors are a fixed point
    U, fit, dr, st = O.cp_als(X, dims, Ftrue, 2, direct=True)
    assert st == 0 and dr[-1] < 1e-6
    # monotone non-increasing error, identity == direct residual
    U0 = [rng.random((d, R)) for d in dims]
    Xn = (X + 0.01 * rng.standard_normal(X.size).astype(np.float32) * np.linalg.norm(X) / np.sqrt(X.size))
    U, fit, dr, st = O.cp_als(Xn.astype(np.float32), dims, U0, 300, direct=True)
    assert np.all(np.diff(dr) <= 1e-12)
    assert np.allclose(fit, dr, rtol=1e-7, atol=1e-9)
    # noise floor (||E|| = 0.01 ||X||) reached
    assert dr[-1] < 0.011


def test_cp_als_order4_rank_recovery():
    rng = np.random.default_rng(9)
    dims = (6, 7, 5, 4)
    R = 2
    Ftrue = [rng.random((d, R)) for d in dims]
    X = O.cp_reconstruct(dims, Ftrue).astype(np.float32)
    U0 = [rng.random((d, R)) for d in dims]
    U, fit, dr, st = O.cp_als(X, dims, U0, 200, direct=True)
    assert dr[-1] < 1e-5


# ---------------------------------------------------------------- Tucker
def test_tucker_golden_and_hosvd():
    g = GOLD["hand_tucker_3x4x2"]
    dims, ranks = g["dims"], g["ranks"]
    U = [np.asarray(g[n], float).reshape((d, r), order="F") for n, d, r in zip("ABC", dims, ranks)]
    Xt = O.tucker_reconstruct(dims, ranks, np.asarray(g["G"], float), U)
    assert np.array_equal(Xt, np.asarray(g["X"]))
    # exact multilinear rank: HOSVD reproduces X, U orthonormal, ||core|| = ||X||
    rng = np.random.default_rng(10)
    dims, ranks = (8, 9, 7), (3, 2, 4)
    core = rng.standard_normal(int(np.prod(ranks)))
    Us = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d, r in zip(dims, ranks)]
    X = O.tucker_reconstruct(dims, ranks, core, Us).astype(np.float32)
    Uh, ch = O.tucker_hosvd(X, dims, ranks)
    for u in Uh:
        assert np.allclose(u.T @ u, np.eye(u.shape[1]), atol=1e-10)
    assert np.isclose(np.linalg.norm(ch), np.linalg.norm(X.astype(np.float64)), rtol=1e-6)
    Xr = O.tucker_reconstruct(dims, ranks, ch, Uh)
    assert np.linalg.norm(Xr - X) / np.linalg.norm(X) < 1e-6
    # the column order is ascending eigenvalue (reference copies the last r syevdx columns)
    X3 = X.reshape(dims, order="F").astype(np.float64)
    G0 = X3.reshape(dims[0], -1, order="F") @ X3.reshape(dims[0], -1, order="F").T
    lam = [Uh[0][:, c] @ G0 @ Uh[0][:, c] for c in range(ranks[0])]
    assert np.all(np.diff(lam) >= -1e-9)
    # HOOI never increases the error
    Xn = (X + 0.05 * rng.standard_normal(X.size).astype(np.float32)).astype(np.float32)
    U0, _ = O.tucker_hosvd(Xn, dims, ranks)
    _, _, hist = O.tucker_hooi(Xn, dims, ranks, 5, U0)
    assert np.all(np.diff(hist) <= 1e-10)


# ---------------------------------------------------------------- MLP
def _mlp_params(rng, nin, width, nout):
    return [rng.standard_normal((width, nin)) * 0.3, rng.standard_normal(width) * 0.1,
            rng.standard_normal((width, width)) * 0.3, rng.standard_normal(width) * 0.1,
            rng.standard_normal((nout, width)) * 0.3, rng.standard_normal(nout) * 0.1]


def test_mlp_gradients_finite_difference():
    rng = np.random.default_rng(11)
    nin, width, nout, batch = 5, 7, 2, 4
    P = _mlp_params(rng, nin, width, nout)
    x = rng.standard_normal((batch, nin))
    t = rng.standard_normal((batch, nout))
    loss, g = O.mlp_backward(x, t, P, nin, width, nout)
    y = O.mlp_forward(x, P, nin, width, nout)
    assert np.isclose(loss, 0.5 * np.mean((y - t) ** 2))
    eps = 1e-6
    for pi, p in enumerate(P):
        flat = p.ravel()
        for idx in range(0, flat.size, max(1, flat.size // 6)):
            old = flat[idx]
            flat[idx] = old + eps
            lp = O.mlp_backward(x, t, P, nin, width, nout)[0]
            flat[idx] = old - eps
            lm = O.mlp_backward(x, t, P, nin, width, nout)[0]
            flat[idx] = old
            assert np.isclose(g[pi].ravel()[idx], (lp - lm) / (2 * eps), rtol=1e-5, atol=1e-8)


def test_adam_matches_paper():
    rng = np.random.default_rng(12)
    p = rng.standard_normal(10)
    m = np.zeros(10)
    v = np.zeros(10)
    p_ref, m_ref, v_ref = p.copy(), m.copy(), v.copy()
    for t in range(1, 4):
        g = rng.standard_normal(10)
        O.adam(p, g, m, v, t, 1e-2)
        m_ref = 0.9 * m_ref + 0.1 * g
        v_ref = 0.999 * v_ref + 0.001 * g * g
        p_ref = p_ref - 1e-2 * (m_ref / (1 - 0.9 ** t)) / (np.sqrt(v_ref / (1 - 0.999 ** t)) + 1e-8)
    assert np.allclose(p, p_ref, rtol=1e-13)


# ---------------------------------------------------------------- MLP known answers
# GELU(x) = x Phi(x) (Hendrycks & Gimpel 2016) and its derivative Phi(x) + x phi(x), computed
# independently at 40 digits with mpmath:  x * (1 + mpmath.erf(x / mpmath.sqrt(2))) / 2
GELU_KAT = {0.0: 0.0, 1.0: 0.84134474606854294859, -1.0: -0.15865525393145705141,
            2.0: 1.9544997361036415856, -3.0: -0.00404969409489028358, 0.5: 0.34573123063700655182}
GELU_D_KAT = {0.0: 0.5, 1.0: 1.0833154705876862984, -1.0: -0.083315470587686298383,
              2.0: 1.0852318010781968967, -3.0: -0.011945647204183927, 0.5: 0.86749512465616284253}
GELU_GELU_KAT = {1.0: 0.67301066416940404567, -1.0: -0.069327625476641394937, 2.0: 1.9050097055649646399}


def _scalar_mlp(w1, b1, w2, b2, w3=1.0, b3=0.0):
    return [np.array([[w1]]), np.array([b1]), np.array([[w2]]), np.array([b2]), np.array([[w3]]), np.array([b3])]


@pytest.mark.parametrize("x", sorted(GELU_KAT))
def test_mlp_gelu_known_answers(x):
    """Width-1 networks isolate one GELU: with W2 = 0 and b2 = x the output is GELU(x); with
    unit weights it is GELU(GELU(x)).  A wrong constant, sign or the tanh approximation fails."""
    y = O.mlp_forward(np.zeros((1, 1)), _scalar_mlp(0.0, 0.0, 0.0, x), 1, 1, 1)
    assert abs(y[0, 0] - GELU_KAT[x]) <= 1e-15 * max(1.0, abs(GELU_KAT[x]))
    if x in GELU_GELU_KAT:
        y = O.mlp_forward(np.array([[x]]), _scalar_mlp(1.0, 0.0, 1.0, 0.0), 1, 1, 1)
        assert abs(y[0, 0] - GELU_GELU_KAT[x]) <= 1e-15


@pytest.mark.parametrize("x", sorted(GELU_D_KAT))
def test_mlp_gelu_derivative_known_answers(x):
    """loss = (y - 0)^2 / 2 with y = GELU(b2): dloss/db2 = GELU(x) GELU'(x)."""
    loss, g = O.mlp_backward(np.zeros((1, 1)), np.zeros((1, 1)), _scalar_mlp(0.0, 0.0, 0.0, x), 1, 1, 1)
    assert abs(loss - 0.5 * GELU_KAT[x] ** 2) <= 1e-15
    assert abs(g[3][0] - GELU_KAT[x] * GELU_D_KAT[x]) <= 1e-15


def test_adam_two_step_known_answer():
    """Kingma & Ba (2015), Algorithm 1, lr 0.1, beta 0.9/0.999, eps 1e-8, p0 = 1, gradients 0.5
    then -0.25; expected values computed at 40 digits with mpmath (m2 = 0.02, v2 = 3.1225e-4)."""
    p, m, v = np.array([1.0]), np.zeros(1), np.zeros(1)
    O.adam(p, np.array([0.5]), m, v, 1, 0.1)
    assert abs(p[0] - 0.90000000199999996) <= 1e-15
    assert abs(m[0] - 0.05) <= 1e-16 and abs(v[0] - 2.5e-4) <= 1e-18
    O.adam(p, np.array([-0.25]), m, v, 2, 0.1)
    assert abs(p[0] - 0.87336629870784616256) <= 1e-15
    assert abs(m[0] - 0.02) <= 1e-16 and abs(v[0] - 3.1225e-4) <= 1e-18


def test_adam_first_step_is_lr_sign():
    """Bias correction makes the first step lr g / (|g| + eps) for every g (m1/c1 = g, v1/c2 = g^2)."""
    g = np.array([3.0, -2e-3, 0.75, -40.0, 1e-6])
    p = np.zeros(5)
    O.adam(p, g, np.zeros(5), np.zeros(5), 1, 0.01)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-14, atol=0)


def test_cp_als3_gemm_matches_loop_oracle():
    """The GEMM formulation (decomposition/CPU/python/mycp.py's MTTKRP as unfolding x Khatri-Rao)
    gives the same FP64 iterates as orc_cp_als; its FP32 variant stays within 1e-5."""
    dims, R = (40, 36, 30), 5
    F = [O.uniform(d * R, 0, m).reshape((d, R), order="F").astype(np.float64) for m, d in enumerate(dims)]
    X0 = O.cp_reconstruct(dims, F)
    E = O.normal(X0.size, 2, 0).astype(np.float64)
    X = (X0 + 0.01 * np.linalg.norm(X0) / np.linalg.norm(E) * E).astype(np.float32)
    U0 = [O.uniform(d * R, 1, m).reshape((d, R), order="F").astype(np.float64) for m, d in enumerate(dims)]
    Uc, fc, dc, st = O.cp_als(X, dims, U0, 15, direct=True)
    Ug, fg, _ = O.cp_als3_gemm(X, dims, U0, 15)
    assert st == 0
    assert np.max(np.abs(dc - fg)) < 1e-12
    for a, b in zip(Uc, Ug):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
    _, f32, _ = O.cp_als3_gemm(X, dims, U0, 15, dtype=np.float32)
    assert np.max(np.abs(f32 - dc)) < 1e-5
    _, fs, _ = O.cp_als3_gemm(X, dims, U0, 15, stop_at=fg[4])
    assert len(fs) == 5


@pytest.mark.parametrize("dims,ranks", [((20, 16, 12), (4, 3, 5)), ((12, 10, 8, 6), (3, 4, 2, 3))])
def test_tucker_hosvd_gemm_matches_loop_oracle(dims, ranks):
    """The GEMM/LAPACK HOSVD used at 128^4 gives the loop oracle's factors (same ascending order
    and sign convention) and core, orders 3 and 4."""
    rng = np.random.default_rng(5)
    core = rng.standard_normal(int(np.prod(ranks)))
    Us = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d, r in zip(dims, ranks)]
    X = O.tucker_reconstruct(dims, ranks, core, Us)
    E = rng.standard_normal(X.size)
    X = (X + 1e-2 * np.linalg.norm(X) / np.linalg.norm(E) * E).astype(np.float32)
    Uo, co = O.tucker_hosvd(X, dims, ranks)
    Ug, cg, w = O.tucker_hosvd_gemm(X, dims, ranks)
    for a, b in zip(Uo, Ug):
        assert np.max(np.abs(a - b)) < 1e-12
    assert np.max(np.abs(co - cg)) < 1e-12 * np.max(np.abs(co))
    assert all(np.all(np.diff(x) >= 0) for x in w)


@pytest.mark.parametrize("dims,ranks", [((20, 16, 12), (4, 3, 5)), ((12, 10, 8, 6), (3, 4, 2, 3))])
def test_tucker_hooi_gemm_matches_loop_oracle(dims, ranks):
    """The GEMM/LAPACK HOOI of the CPU baseline gives the loop oracle's factors, core and
    per-iteration errors (orders 3 and 4, 3 iterations from the HOSVD factors)."""
    rng = np.random.default_rng(6)
    core = rng.standard_normal(int(np.prod(ranks)))
    Us = [np.linalg.qr(rng.standard_normal((d, r)))[0] for d, r in zip(dims, ranks)]
    X = O.tucker_reconstruct(dims, ranks, core, Us)
    E = rng.standard_normal(X.size)
    X = (X + 5e-2 * np.linalg.norm(X) / np.linalg.norm(E) * E).astype(np.float32)
    U0, _ = O.tucker_hosvd(X, dims, ranks)
    Uo, co, ho = O.tucker_hooi(X, dims, ranks, 3, U0)
    Ug, cg, hg, _ = O.tucker_hooi_gemm(X, dims, ranks, 3, U0)
    for a, b in zip(Uo, Ug):
        assert np.max(np.abs(a - b)) < 1e-10
    assert np.max(np.abs(co - cg)) < 1e-10 * np.max(np.abs(co))
    assert np.max(np.abs(ho - hg)) < 1e-12


@pytest.mark.parametrize("dims,R", [((10, 9, 8, 7), 4), ((14, 12, 10), 5)])
def test_cp_als_gemm_matches_loop_oracle(dims, R):
    """The order-N GEMM ALS of the CPU baseline (C5 is order 4) gives the loop oracle's
    per-sweep fast-fit errors and factors (orders 4 and 3)."""
    rng = np.random.default_rng(7)
    F = [rng.standard_normal((d, R)) for d in dims]
    X = O.cp_reconstruct(dims, F)
    E = rng.standard_normal(X.size)
    X = (X + 1e-2 * np.linalg.norm(X) / np.linalg.norm(E) * E).astype(np.float32)
    U0 = [rng.uniform(size=(d, R)) for d in dims]
    Uo, fo, _, _ = O.cp_als(X, dims, U0, 12)
    Ug, fg, _ = O.cp_als_gemm(X, dims, U0, 12)
    assert np.max(np.abs(fo - fg)) < 1e-10
    for a, b in zip(Uo, Ug):
        assert np.max(np.abs(a - b)) < 1e-8 * np.max(np.abs(a))
    _, fs, _ = O.cp_als_gemm(X, dims, U0, 12, stop_at=fg[3])
    assert len(fs) == 4

