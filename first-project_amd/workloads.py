"""Synthetic inputs of the five benchmark recipes, generated on the device with the
library's Philox streams (identical to the oracle's generator).

  C1  128^3, rank 16, U(0,1) factors (seed 0), 1% noise, U(0,1) init (seed 1)
  C2  1024^3, rank 64, N(0,1) unit-column factors (seed 0), 0.1% noise, N(0,1) init (seed 1)
  C3  128^4 Tucker, 32^4 N(0,1) core (seed 0), orthonormal factors from QR of N(0,1) 128x32
      blocks (seeds 1..4), 0.5% noise (seed 5)
  C4  4096 tensors, extents U{16..96}, ranks {4,8,16}, U(0,1) factors (seed t), 1% noise
  C5  96^4, rank 24, factor columns with pairwise congruence 0.9, 0.1% noise
      (extrapolation.swamp_tensor)
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import tdx  # noqa: E402


def _uniform(n, seed, stream):
    import torch
    t = torch.empty(n, device="cuda")
    tdx.fill_uniform(t, seed, stream)
    return t


def _normal(n, seed, stream):
    import torch
    t = torch.empty(n, device="cuda")
    tdx.fill_normal(t, seed, stream)
    return t


def unit_columns(flat, d, R):
    M = flat.view(R, d).t()
    M = M / M.norm(dim=0, keepdim=True)
    return M.t().contiguous().view(-1)


def cp_tensor(dims, R, kind, eta, factor_seed=0, noise_seed=2):
    import torch
    if kind == "uniform":
        F = [_uniform(d * R, factor_seed, m) for m, d in enumerate(dims)]
    else:
        F = [unit_columns(_normal(d * R, factor_seed, m), d, R) for m, d in enumerate(dims)]
    X = torch.empty(int(np.prod(dims)), device="cuda")
    tdx.gen_cp_tensor(X, dims, F, R, eta, noise_seed, 0)
    return X, F


def uniform_init(dims, R, seed=1):
    return [_uniform(d * R, seed, m) for m, d in enumerate(dims)]


def normal_init(dims, R, seed=1):
    return [_normal(d * R, seed, m) for m, d in enumerate(dims)]


def c1():
    dims, R = (128, 128, 128), 16
    X, F = cp_tensor(dims, R, "uniform", 1e-2)
    return X, dims, R


def c2(n=1024, R=64):
    dims = (n, n, n)
    X, F = cp_tensor(dims, R, "normal", 1e-3)
    return X, dims, R


def c3(n=128, r=32, order=4, eta=5e-3, with_factors=False):
    """Tucker tensor core x_1 U_1 ... x_d U_d + noise (noise scaled to eta ||X0||); with
    with_factors=True also the generating orthonormal factors (host FP64, n x r)."""
    import torch
    dims = (n,) * order
    ranks = (r,) * order
    core = _normal(r ** order, 0, 0)
    Us, Qs = [], []
    for m in range(order):
        G = _normal(n * r, m + 1, 0).cpu().numpy().reshape((n, r), order="F").astype(np.float64)
        Q, _ = np.linalg.qr(G)
        Qs.append(Q)
        Us.append(torch.from_numpy(np.ascontiguousarray(Q.ravel(order="F"), np.float32)).cuda())
    X0 = torch.empty(n ** order, device="cuda")
    if order == 4:
        tdx.gentucker4(X0, core, *Us, *dims, *ranks)
    else:
        tdx.gentucker(X0, core, *Us, *dims, *ranks)
    E = _normal(n ** order, 5, 0)
    X = X0 + (eta * X0.double().norm() / E.double().norm()).float() * E
    if with_factors:
        return X.contiguous(), dims, ranks, Qs
    return X.contiguous(), dims, ranks


def c4(ntensors=4096, seed=0, lo=16, hi=96):
    """Packed batch of uneven tensors; extents and ranks from a seeded numpy generator."""
    import torch
    rng = np.random.default_rng(seed)
    shapes = rng.integers(lo, hi + 1, size=(ntensors, 3))
    ranks = rng.choice([4, 8, 16], size=ntensors)
    b = tdx.Batch(shapes, ranks)
    X = torch.empty(b.x_floats, device="cuda")
    for t in range(ntensors):
        dims = tuple(int(v) for v in b.shapes[t])
        R = int(b.ranks[t])
        F = [_uniform(d * R, 10_000 + t, m) for m, d in enumerate(dims)]
        xt = X[int(b.xoff[t]):int(b.xoff[t]) + int(np.prod(dims))]
        tdx.gen_cp_tensor(xt, dims, F, R, 1e-2, 20_000 + t, 0)
    return X, b


def run_c3(iters=10, bf16=False, history=False, data=None):
    """HOSVD + `iters` HOOI iterations of the C3 tensor (BF16 Grams and first TTMs with
    bf16=True); returns (seconds, final relative error from the core, per-iteration errors or
    None).  The final error is sqrt(1 - ||core||^2/||X||^2) in FP64, outside the timed call."""
    import time

    import torch
    X, dims, ranks = data if data is not None else c3()
    U = [torch.empty(d * r, device="cuda") for d, r in zip(dims, ranks)]
    core = torch.empty(int(np.prod(ranks)), device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = tdx.tucker_hooi(X, dims, ranks, iters, core, U, bf16_ttm=bf16, history=history)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    nx2 = float(X.double().square().sum())
    nc2 = float(core.double().square().sum())
    return sec, float(np.sqrt(max(nx2 - nc2, 0.0) / nx2)), hist


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="run one benchmark recipe on cuda:0")
    ap.add_argument("recipe", choices=["c3", "c4"])
    ap.add_argument("--iters", type=int, default=10, help="c3: HOOI iterations")
    ap.add_argument("--bf16", action="store_true", help="c3: BF16 Grams and first TTMs")
    ap.add_argument("--history", action="store_true", help="c3: relative error after every iteration")
    ap.add_argument("--sweeps", type=int, default=3, help="c4: batched ALS sweeps")
    ap.add_argument("--variant", type=int, default=-1, help="c4: batched kernel variant (-1: default)")
    a = ap.parse_args()
    tdx.lib()
    if a.recipe == "c3":
        sec, rel, hist = run_c3(a.iters, a.bf16, a.history)
        print(f"c3 ({'bf16' if a.bf16 else 'fp32'}): {sec:.4f} s, final rel-err {rel:.6e}"
              + (f", per iteration {list(map(float, hist))}" if hist is not None else ""))
    else:
        import torch
        X, b = c4()
        U = torch.empty(b.u_floats, device="cuda")
        st, fit, ms = tdx.cp_als_batched(X, U, b, a.sweeps, init=0, seed=1, variant=a.variant)
        gbs = 3 * 4 * b.x_floats / (float(np.mean(ms)) * 1e-3) / 1e9
        print(f"c4: {b.n} tensors, {b.x_floats} elements, variant {a.variant}: ms per sweep "
              f"{[round(float(v), 3) for v in ms]} ({gbs:.0f} GB/s of 3 passes over X), median final relerr "
              f"{float(np.median(fit[:, -1])):.5f}")
