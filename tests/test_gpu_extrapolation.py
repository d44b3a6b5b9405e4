"""The extrapolation step inside the ALS sweep graph: the candidate evaluation is exact, the
accept rule never lets the error grow, the grid mode leaves the plain trajectory unchanged,
and a small trained MLP reaches the target in fewer sweeps than plain ALS on swamp tensors
(reduced 32^4 rank-6 recipe so the test runs in seconds)."""
import numpy as np
import pytest

import extrapolation as E
import tdx

pytestmark = pytest.mark.gpu

DIMS, R = (32, 32, 32, 32), 6


def test_grid_mode_is_plain_and_alpha_one_is_exact():
    X = E.swamp_tensor(3, DIMS, R)
    U = E.init_factors(3, DIMS, R)
    fit_p, _, _ = E.run(X, DIMS, R, U, tdx.LS_NONE, 12)
    U = E.init_factors(3, DIMS, R)
    fit_g, _, tr = E.run(X, DIMS, R, U, tdx.LS_GRID, 12, grid=[1.0, 1.5, 3.0])
    assert np.array_equal(fit_p, fit_g)
    assert np.array_equal(tr["plain_relerr"], fit_p)
    # alpha = 1 reproduces the plain update up to the FP32 rounding of U1 + 1*(U - U1) and of
    # the two different MTTKRPs behind the two fit identities (largest at sweep 0, where the
    # swamp factors have large cancelling columns)
    assert np.max(np.abs(tr["grid_relerr"][:, 0] - fit_p)) < 1e-5


def test_lsb_accept_rule_and_exact_evaluation():
    X = E.swamp_tensor(4, DIMS, R)
    U = E.init_factors(4, DIMS, R)
    fit, ms, tr = E.run(X, DIMS, R, U, tdx.LS_LSB, 30)
    assert np.all(np.diff(fit) <= 1e-7)
    acc = tr["accepted"].astype(bool)
    assert acc.any()
    assert np.all(fit[acc] < tr["plain_relerr"][acc])
    assert np.all(fit[~acc] == tr["plain_relerr"][~acc])
    assert np.allclose(tr["alpha"], np.cbrt(np.arange(1, 31)), rtol=1e-6)
    # the kept factors really have the recorded error (direct compensated residual)
    direct = tdx.cp_relerr(X, DIMS, U, R)
    assert abs(direct - fit[-1]) < 2e-6


@pytest.mark.parametrize("storage", [tdx.TDX_BF16, tdx.TDX_FP8_E4M3])
def test_line_search_with_reduced_storage_is_rejected(storage):
    # switched off for BF16 / FP8 storage: the call returns before any device work
    X = E.swamp_tensor(5, DIMS, R)
    U = E.init_factors(5, DIMS, R)
    before = [u.clone() for u in U]
    o = tdx.default_cp_options()
    o.max_sweeps, o.refine_sweeps, o.storage, o.init, o.line_search = 4, 1, storage, 1, tdx.LS_LSB
    with pytest.raises(tdx.TdxError) as e:
        tdx.cp_als_ex(X, DIMS, R, U, o)
    assert e.value.status == tdx.TDX_ERR_UNSUPPORTED
    assert all(bool((a == b).all()) for a, b in zip(U, before))


def test_learned_step_gains_over_plain():
    feats, labels = E.collect(range(24), DIMS, R, sweeps=40)
    assert feats.shape == (24 * 40, tdx.LS_FEATURES)
    assert np.all(np.isfinite(feats)) and np.all(labels >= 0)
    P, hist = E.train(feats, labels, epochs=80, batch=128, lr=3e-3)
    assert hist[-1] < hist[0]
    rows, target = E.evaluate(range(2000, 2006), P, DIMS, R, max_sweeps=300)
    s = E.summarize(rows)
    print("extrapolation summary:", s)
    assert s["learned"]["reached"] >= s["plain"]["reached"]
    assert s["learned"]["median_sweeps"] < s["plain"]["median_sweeps"]
