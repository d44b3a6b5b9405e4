"""GPU parity of the fused MTTKRP against the FP64 oracle.

Bars: bit-exact where the arithmetic is exact (small-integer data, as in the reference's
test/test_mttkrp.cpp:15-27), FP32 storage within 1e-5 normwise of the FP64 oracle, BF16
and FP8 storage within 3e-3 normwise of the oracle evaluated on the same stored values
(the remaining difference is the BF16 rounding of the on-the-fly Khatri-Rao rows).
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


def factors(rng, dims, R, kind="normal"):
    if kind == "int":
        return [rng.integers(-4, 5, size=(d, R)).astype(np.float32) for d in dims]
    return [rng.standard_normal((d, R)).astype(np.float32) for d in dims]


def run_mttkrp(Xdev, dtype, scales, dims, U, mode):
    R = U[0].shape[1]
    Ud = [dev(u.ravel(order="F")) for u in U]
    out = empty(dims[mode] * R)
    tdx.mttkrp_ex(Xdev, dtype, scales, dims, Ud, R, mode, out)
    return host(out).reshape((dims[mode], R), order="F")


def test_golden_srand2_all_modes_exact():
    for key in ("mttkrp_srand2", "hand_cp_3x4x2"):
        g = GOLD[key]
        m, n, k = g["dims"]
        r = g["rank"]
        X, A, B, C = [dev(np.asarray(g[x], np.float32)) for x in "XABC"]
        for flag, (left, right) in zip((1, 2, 3), ((C, B), (C, A), (B, A))):
            out = empty(g["dims"][flag - 1] * r)
            tdx.mttkrp(X, left, right, out, flag, m, n, k, r)
            assert np.array_equal(host(out), np.asarray(g["M"][flag - 1], np.float32)), (key, flag)


# (48, 40): order 2, so output mode 1 runs the register-tile kernel, not the LDS-staged one
@pytest.mark.parametrize("dims,R", [((64, 32, 48), 16), ((40, 24, 64), 8), ((128, 64, 32), 64), ((16, 40, 20, 8), 4),
                                    ((48, 40), 16)])
@pytest.mark.parametrize("dtype", [tdx.TDX_F32, tdx.TDX_BF16])
def test_integer_data_bit_exact(dims, R, dtype):
    import torch
    rng = np.random.default_rng(sum(dims) + R)
    X = rng.integers(-4, 5, size=int(np.prod(dims))).astype(np.float32)
    U = factors(rng, dims, R, "int")
    Xd = dev(X)
    if dtype == tdx.TDX_BF16:
        q = torch.empty(X.size, dtype=torch.int16, device="cuda")
        tdx.quantize_bf16(Xd, q)
        Xd = q
    for mode in range(len(dims)):
        got = run_mttkrp(Xd, dtype, None, dims, U, mode)
        ref = O.mttkrp(X, dims, [u.astype(np.float64) for u in U], mode)
        assert np.array_equal(got, ref.astype(np.float32)), (dims, mode, dtype)


SHAPES = [((128, 128, 128), 16), ((96, 64, 80), 24), ((32, 96, 40), 64), ((24, 40, 33), 5),
          ((48, 40, 32, 24), 24), ((5, 7, 9), 3), ((16, 12, 10), 80), ((256, 48, 64), 48)]


@pytest.mark.parametrize("dims,R", SHAPES)
def test_fp32_within_1e5_of_fp64(dims, R):
    rng = np.random.default_rng(len(dims) * 100 + R)
    X = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    U = factors(rng, dims, R)
    Xd = dev(X)
    for mode in range(len(dims)):
        got = run_mttkrp(Xd, tdx.TDX_F32, None, dims, U, mode)
        ref = O.mttkrp(X, dims, [u.astype(np.float64) for u in U], mode)
        err = relnorm(got, ref)
        assert err <= 1e-5, (dims, R, mode, err)


@pytest.mark.parametrize("dims,R", [s for s in SHAPES if s[0][0] % 8 == 0 or s[1] > 64])
def test_bf16_storage(dims, R):
    import torch
    rng = np.random.default_rng(len(dims) * 10 + R)
    X = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    U = factors(rng, dims, R)
    q = torch.empty(X.size, dtype=torch.int16, device="cuda")
    tdx.quantize_bf16(dev(X), q)
    Xb = (O.bf16_bits(X).astype(np.uint32) << 16).view(np.float32)
    for mode in range(len(dims)):
        got = run_mttkrp(q, tdx.TDX_BF16, None, dims, U, mode)
        ref = O.mttkrp(Xb, dims, [u.astype(np.float64) for u in U], mode)
        err = relnorm(got, ref)
        assert err <= 3e-3, (dims, R, mode, err)


# (24, 40, 32): I_0 = 24 holds 8-element but not 16-element FP8 vectors, so output modes 1-2
# run the register-tile kernel instead of the LDS-staged one
@pytest.mark.parametrize("dims,R", [((128, 128, 128), 16), ((96, 64, 80), 24), ((64, 40, 24, 16), 16),
                                    ((16, 12, 10), 80), ((24, 40, 32), 32)])
def test_fp8_storage(dims, R):
    import torch
    rng = np.random.default_rng(7 + R)
    X = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    U = factors(rng, dims, R)
    I = dims[0]
    q = torch.empty(X.size, dtype=torch.uint8, device="cuda")
    s = empty(I)
    tdx.quantize_fp8_mode1(dev(X), q, s, I)
    Qo, so = O.quantize_e4m3_mode1(X, I)
    dec = O.e4m3_decode(Qo)
    for mode in range(len(dims)):
        got = run_mttkrp(q, tdx.TDX_FP8_E4M3, s, dims, U, mode)
        ref = O.mttkrp(dec, dims, [u.astype(np.float64) for u in U], mode, scales=so)
        err = relnorm(got, ref)
        assert err <= 3e-3, (dims, R, mode, err)
    # and the FP8 result is a faithful MTTKRP of X itself (quantisation error only)
    ref32 = O.mttkrp(X, dims, [u.astype(np.float64) for u in U], 0)
    assert relnorm(run_mttkrp(q, tdx.TDX_FP8_E4M3, s, dims, U, 0), ref32) < 0.05


@pytest.mark.parametrize("I", [8, 16])
def test_fp8_all_codes_decode(I):
    """Every finite E4M3 code through the MTTKRP: X = codes (scale 1 per slice would need
    amax 448), factors of ones, so mode-0 output row i = sum of its decoded codes.  I = 8
    runs the register-tile kernel, I = 16 the LDS kernel's FP8 tile (transposing LDS reads)."""
    import torch
    codes = np.array([c for c in range(256) if (c & 0x7f) != 0x7f], np.uint8)  # 254 finite codes
    J = 256 // I
    X = np.zeros(I * J, np.uint8)
    X[:codes.size] = codes
    q = dev(X).to(torch.uint8)
    s = dev(np.ones(I, np.float32))
    dims = (I, J, 1)
    U = [np.ones((I, 16), np.float32), np.zeros((J, 16), np.float32), np.ones((1, 16), np.float32)]
    for j in range(J):
        U1 = np.zeros((J, 16), np.float32)
        U1[j, :] = 1.0
        U[1] = U1
        got = run_mttkrp(q, tdx.TDX_FP8_E4M3, s, dims, U, 0)[:, 0]
        assert np.array_equal(got, O.e4m3_decode(X[j * I:(j + 1) * I])), j


@pytest.mark.parametrize("dims,R", [((256, 96, 3), 16), ((128, 200, 2), 64), ((48, 130, 5), 24),
                                    ((144, 72, 4, 3), 32)])
def test_fp8_mode0_exact(dims, R):
    """Mode 0 from FP8 storage (the LDS kernel's FP8 tile: full and partial row tiles, full
    and partial i_1 chunks, order 4) on codes of {0, +-0.5, +-1, +-2} with slice scales 1 and
    integer factors: every product and partial sum is exact in FP32, so the result equals the
    oracle's bit for bit."""
    import torch
    dec_all = O.e4m3_decode(np.arange(256, dtype=np.uint8))
    pick = [int(np.flatnonzero(dec_all == v)[0]) for v in (0.0, 0.5, -0.5, 1.0, -1.0, 2.0, -2.0)]
    rng = np.random.default_rng(sum(dims) + R)
    Xc = np.array(pick, np.uint8)[rng.integers(0, len(pick), size=int(np.prod(dims)))]
    U = [rng.integers(-2, 3, size=(d, R)).astype(np.float32) for d in dims]
    q = dev(Xc).to(torch.uint8)
    s = dev(np.ones(dims[0], np.float32))
    got = run_mttkrp(q, tdx.TDX_FP8_E4M3, s, dims, U, 0)
    ref = O.mttkrp(O.e4m3_decode(Xc), dims, [u.astype(np.float64) for u in U], 0)
    assert np.array_equal(got, ref.astype(np.float32))


# ---------------------------------------------------------------------------
# shared pass of an order-3 sweep: M1 and M2 from T = X x_0 A^T

def run_dtree(Xdev, dtype, scales, dims, U):
    R = U[0].shape[1]
    Ud = [dev(u.ravel(order="F")) for u in U]
    M1, M2 = empty(dims[1] * R), empty(dims[2] * R)
    tdx.mttkrp_dtree(Xdev, dtype, scales, dims, Ud, R, M1, M2)
    return (host(M1).reshape((dims[1], R), order="F"), host(M2).reshape((dims[2], R), order="F"))


@pytest.mark.parametrize("dims,R", [((256, 64, 40), 16), ((256, 160, 9), 64), ((512, 32, 70), 24)])
def test_shared_pass_integer_data_bit_exact(dims, R):
    """Small integers: BF16 storage and BF16 A operands are exact and every partial sum is an
    integer below 2^24, so M1 equals the oracle's MTTKRP bit for bit, and M2 equals, bit for
    bit, the contraction of T = X x_0 A^T rounded to BF16 (to nearest even: T is stored in BF16)
    with B; on these inputs T exceeds 256 in magnitude, so the rounding is exercised."""
    import torch
    rng = np.random.default_rng(sum(dims) + 3 * R)
    X = rng.integers(-3, 4, size=int(np.prod(dims))).astype(np.float32)
    U = factors(rng, dims, R, "int")
    q = torch.empty(X.size, dtype=torch.int16, device="cuda")
    tdx.quantize_bf16(dev(X), q)
    M1, M2 = run_dtree(q, tdx.TDX_BF16, None, dims, U)
    U64 = [u.astype(np.float64) for u in U]
    assert np.array_equal(M1, O.mttkrp(X, dims, U64, 1).astype(np.float32))
    X3 = X.reshape(dims, order="F").astype(np.float64)
    T = np.einsum("abc,ar->bcr", X3, U64[0])  # exact integers below 2^24
    assert np.max(np.abs(T)) > 256
    Tb = (O.bf16_bits(T.astype(np.float32)).astype(np.uint32) << 16).view(np.float32).reshape(T.shape)
    M2_ref = np.einsum("bcr,br->cr", Tb.astype(np.float64), U64[1])
    assert np.array_equal(M2, M2_ref.astype(np.float32))
    # and within BF16 rounding of the unrounded MTTKRP
    assert relnorm(M2, O.mttkrp(X, dims, U64, 2)) <= 3e-3


@pytest.mark.parametrize("dims,R", [((256, 96, 48), 16), ((512, 64, 40), 32), ((256, 128, 24), 64), ((256, 32, 5), 10),
                                    ((2048, 32, 6), 32), ((256, 64, 40), 48), ((512, 96, 17), 40)])
def test_shared_pass_bf16_storage(dims, R):
    """Within 3e-3 normwise of the oracle on the stored values (BF16 rounding of A and, for M2,
    of T), and within 6e-3 of the fused per-mode kernel, which meets the same 3e-3 bar with its
    own rounding (of the Khatri-Rao block): the two bars add."""
    import torch
    rng = np.random.default_rng(len(dims) * 17 + R)
    X = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    U = factors(rng, dims, R)
    q = torch.empty(X.size, dtype=torch.int16, device="cuda")
    tdx.quantize_bf16(dev(X), q)
    Xb = (O.bf16_bits(X).astype(np.uint32) << 16).view(np.float32)
    M1, M2 = run_dtree(q, tdx.TDX_BF16, None, dims, U)
    U64 = [u.astype(np.float64) for u in U]
    for mode, got in ((1, M1), (2, M2)):
        ref = O.mttkrp(Xb, dims, U64, mode)
        assert relnorm(got, ref) <= 3e-3, (dims, R, mode, relnorm(got, ref))
        assert relnorm(got, run_mttkrp(q, tdx.TDX_BF16, None, dims, U, mode)) <= 6e-3


@pytest.mark.parametrize("dims,R", [((512, 64, 40), 16), ((512, 96, 33), 64), ((1024, 32, 24), 32)])
def test_shared_pass_fp8_storage(dims, R):
    import torch
    rng = np.random.default_rng(31 + R)
    X = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    U = factors(rng, dims, R)
    I = dims[0]
    q = torch.empty(X.size, dtype=torch.uint8, device="cuda")
    s = empty(I)
    tdx.quantize_fp8_mode1(dev(X), q, s, I)
    Qo, so = O.quantize_e4m3_mode1(X, I)
    dec = O.e4m3_decode(Qo)
    M1, M2 = run_dtree(q, tdx.TDX_FP8_E4M3, s, dims, U)
    U64 = [u.astype(np.float64) for u in U]
    for mode, got in ((1, M1), (2, M2)):
        ref = O.mttkrp(dec, dims, U64, mode, scales=so)
        assert relnorm(got, ref) <= 3e-3, (dims, R, mode, relnorm(got, ref))


@pytest.mark.parametrize("dims,R,dtype", [((128, 64, 40), 16, 1), ((384, 64, 40), 16, 1), ((256, 48, 40), 16, 1),
                                          ((256, 64, 40), 16, 2), ((256, 64, 40), 80, 1)])
def test_shared_pass_rejects_uncovered_shapes(dims, R, dtype):
    """I0 not 128 k (BF16) / 256 k (FP8) with k in {2, 4, 8, 16}, I1 not a multiple of 32 or
    rank > 64: TDX_ERR_UNSUPPORTED, nothing launched."""
    import torch
    dt = tdx.TDX_BF16 if dtype == 1 else tdx.TDX_FP8_E4M3
    n = int(np.prod(dims))
    q = torch.zeros(n, dtype=torch.int16 if dtype == 1 else torch.uint8, device="cuda")
    s = dev(np.ones(dims[0], np.float32)) if dtype == 2 else None
    U = [np.ones((d, R), np.float32) for d in dims]
    with pytest.raises(tdx.TdxError) as e:
        run_dtree(q, dt, s, dims, U)
    assert e.value.status == tdx.TDX_ERR_UNSUPPORTED
