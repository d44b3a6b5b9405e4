"""Batched CP-ALS (one persistent kernel per sweep, or the split design: one MTTKRP and one
solve kernel per mode) against the FP64 oracle and against the single-tensor path, on small
batches of uneven tensors of the batched recipe (extents in [16, 96], ranks in {4, 8, 16},
U(0,1) factors, 1% noise)."""
import numpy as np
import pytest

import oracle as O
import tdx
from helpers import dev, empty, host

pytestmark = pytest.mark.gpu


def make_batch(n, seed=0, lo=16, hi=96):
    import torch
    rng = np.random.default_rng(seed)
    shapes = rng.integers(lo, hi + 1, size=(n, 3))
    ranks = rng.choice([4, 8, 16], size=n)
    b = tdx.Batch(shapes, ranks)
    X = torch.empty(b.x_floats, device="cuda")
    for t in range(n):
        dims = tuple(int(v) for v in b.shapes[t])
        R = int(b.ranks[t])
        F = [torch.from_numpy(O.uniform(d * R, 10_000 + t, m)).cuda() for m, d in enumerate(dims)]
        xt = X[b.xoff[t]:b.xoff[t] + int(np.prod(dims))]
        tdx.gen_cp_tensor(xt, dims, F, R, 1e-2, 20_000 + t, 0)
    return b, X


def test_batched_matches_oracle_per_tensor():
    import torch
    b, X = make_batch(12, seed=1, lo=16, hi=40)
    U = torch.empty(b.u_floats, device="cuda")
    sweeps = 10
    st, fit, ms = tdx.cp_als_batched(X, U, b, sweeps, init=0, seed=77)
    assert st == tdx.TDX_OK
    Xh = host(X)
    for t in range(b.n):
        dims = tuple(int(v) for v in b.shapes[t])
        R = int(b.ranks[t])
        U0 = [O.uniform(d * R, 77 + t, m).reshape((d, R), order="F") for m, d in enumerate(dims)]
        xt = Xh[b.xoff[t]:b.xoff[t] + int(np.prod(dims))]
        _, _, dr, _ = O.cp_als(xt, dims, U0, sweeps, direct=True)
        assert np.max(np.abs(fit[t] - dr)) <= 2e-5, (t, dims, R, np.max(np.abs(fit[t] - dr)))


def test_batched_equals_single_tensor_path_and_variants_agree():
    import torch
    b, X = make_batch(40, seed=2)
    U = torch.empty(b.u_floats, device="cuda")
    st, fit, ms = tdx.cp_als_batched(X, U, b, 6, init=0, seed=5, variant=0)
    for v in range(1, tdx.lib().tdx_cp_batched_variants()):
        U2 = torch.empty(b.u_floats, device="cuda")
        _, fit2, _ = tdx.cp_als_batched(X, U2, b, 6, init=0, seed=5, variant=v)
        assert np.max(np.abs(fit2 - fit)) < 1e-5, v
    # the single-tensor driver from the same init tracks the batched fits
    for t in (0, 7, 23):
        dims = tuple(int(v) for v in b.shapes[t])
        R = int(b.ranks[t])
        U0 = [O.uniform(d * R, 5 + t, m) for m, d in enumerate(dims)]
        Ud = [dev(u) for u in U0]
        o = tdx.default_cp_options()
        o.max_sweeps, o.init = 6, 1
        xt = X[b.xoff[t]:b.xoff[t] + int(np.prod(dims))]
        _, _, f1, _ = tdx.cp_als_ex(xt, dims, R, Ud, o)
        assert np.max(np.abs(f1 - fit[t])) < 1e-5


def test_split_design_equals_persistent_bit_for_bit():
    """The split design (also with mode 2 from the stored T = X x_0 A^T) runs the persistent
    kernel's arithmetic in the same summation orders, so at the same block size and slab the
    factors and fits agree bit for bit.  3000 tensors:
    more than the workgroups resident at once, so every workgroup takes several tensors from
    the queue (the case where a workgroup's factor blocks must be rewritten safely)."""
    import torch
    L = tdx.lib()
    names = [L.tdx_cp_batched_variant_name(v).decode() for v in range(L.tdx_cp_batched_variants())]
    b, X = make_batch(3000, seed=4, lo=16, hi=28)
    for slab in ("4k", "6k"):
        out = []
        for name in (f"bs256_slab{slab}", f"split_bs256_slab{slab}", f"split_t_bs256_slab{slab}"):
            U = torch.empty(b.u_floats, device="cuda")
            _, fit, _ = tdx.cp_als_batched(X, U, b, 2, init=0, seed=9, variant=names.index(name))
            out.append((fit, host(U)))
        for o in out[1:]:
            assert np.array_equal(out[0][0], o[0]), slab
            assert np.array_equal(out[0][1], o[1]), slab


def test_autotuner_picks_and_caches(tmp_path):
    import torch
    b, X = make_batch(64, seed=3)
    U = torch.empty(b.u_floats, device="cuda")
    tdx.cp_als_batched(X, U, b, 0, init=0, seed=1)  # init only
    cache = str(tmp_path / "tune.txt")
    best, times = tdx.cp_batched_autotune(X, U, b, trial_sweeps=2, cache_path=cache)
    assert 0 <= best < len(times)
    assert times[list(times)[best]] == min(times.values())
    best2, _ = tdx.cp_batched_autotune(X, U, b, trial_sweeps=2, cache_path=cache)
    assert best2 == best
    assert open(cache).read().count("\n") == 1
