"""Learned ALS extrapolation: data, training and evaluation on the device.

The step: after a plain ALS sweep U_start -> U, the candidate U_a = U_start + alpha (U - U_start)
is evaluated exactly and kept only if its relative error is lower (the accept rule of
Tensorlab cpd_als.m:296-317; plain ALS stays the fallback).  alpha comes from a 3-layer,
width-64 GELU MLP of the sweep's features (see TDX_LS_FEATURES in include/tdx.h).

Pipeline (all GPU work through libtdx):
  1. `swamp_tensor(seed)`: 96^4 rank-24 tensors whose factor columns have pairwise
     congruence 0.9 (columns = sqrt(c) shared + sqrt(1-c) independent N(0,1), normalised),
     plus 0.1% Gaussian noise -- the recipe that produces ALS swamps.
  2. `collect(seeds)`: plain ALS with the grid mode records, every sweep, the features and the
     exact relative error at each grid alpha; the label is log(alpha*) of the best grid point
     (0 when no grid point beats the plain update).
  3. `train(...)`: Adam on the device (tdx_mlp_train) with FP32 weights.
  4. `evaluate(seeds)`: sweeps and device seconds to reach the target relative error, for
     plain ALS, the cpd_lsb rule alpha = (k+1)^(1/3), and the learned step, all from the same
     initial factors.

Weights are saved as <name>.bin (raw FP32, the libtdx parameter layout) + <name>.json.
"""
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import tdx  # noqa: E402

GRID = [1.25, 1.5, 2.0, 3.0, 4.0, 6.0, 9.0, 14.0]
WIDTH = 64


def swamp_factors(seed, dims, R, congruence=0.9):
    import torch
    F = []
    for m, d in enumerate(dims):
        shared = torch.empty(d, device="cuda")
        indep = torch.empty(d * R, device="cuda")
        tdx.fill_normal(shared, seed, 100 + m)
        tdx.fill_normal(indep, seed, m)
        cols = (congruence ** 0.5) * shared[:, None] + ((1 - congruence) ** 0.5) * indep.view(R, d).t()
        cols = cols / cols.norm(dim=0, keepdim=True)
        F.append(cols.t().contiguous().view(-1))  # column-major d x R
    return F


def swamp_tensor(seed, dims=(96, 96, 96, 96), R=24, eta=1e-3, congruence=0.9, out=None):
    import torch
    F = swamp_factors(seed, dims, R, congruence)
    X = out if out is not None else torch.empty(int(np.prod(dims)), device="cuda")
    tdx.gen_cp_tensor(X, dims, F, R, eta, 1_000_000 + seed, 0)
    return X


def init_factors(seed, dims, R):
    import torch
    U = []
    for m, d in enumerate(dims):
        t = torch.empty(d * R, device="cuda")
        tdx.fill_uniform(t, 2_000_000 + seed, m)
        U.append(t)
    return U


def run(X, dims, R, U, mode, sweeps, params=None, target=0.0, grid=None):
    o = tdx.default_cp_options()
    o.max_sweeps, o.init, o.record_fit, o.use_graph = sweeps, 1, 1, 1
    o.line_search = mode
    o.target_relerr = target
    if mode == tdx.LS_LEARNED:
        o.mlp_params = params.data_ptr()
        o.mlp_width = WIDTH
    if mode == tdx.LS_GRID:
        g = grid or GRID
        o.ngrid = len(g)
        for i, a in enumerate(g):
            o.grid[i] = a
    trace = {} if mode else None
    st, rep, fit, ms = tdx.cp_als_ex(X, dims, R, U, o, want_ms=True, trace=trace)
    return fit, ms, trace


def collect(seeds, dims=(96, 96, 96, 96), R=24, sweeps=60, log=None):
    """(features, labels) from plain-ALS trajectories with the grid evaluation."""
    import torch
    X = torch.empty(int(np.prod(dims)), device="cuda")
    feats, labels = [], []
    for i, sd in enumerate(seeds):
        swamp_tensor(sd, dims, R, out=X)
        U = init_factors(sd, dims, R)
        fit, ms, tr = run(X, dims, R, U, tdx.LS_GRID, sweeps)
        g = tr["grid_relerr"]
        plain = tr["plain_relerr"]
        best = np.argmin(g, axis=1)
        better = g[np.arange(len(best)), best] < plain
        alpha = np.where(better, np.asarray(GRID)[best], 1.0)
        feats.append(tr["features"])
        labels.append(np.log(alpha).astype(np.float32))
        if log and (i + 1) % max(1, len(seeds) // 10) == 0:
            log(f"collect {i + 1}/{len(seeds)}")
    return np.concatenate(feats), np.concatenate(labels)


def train(features, labels, epochs=60, batch=512, lr=2e-3, seed=0):
    import torch
    n = features.shape[0]
    P = torch.empty(tdx.mlp_param_count(tdx.LS_FEATURES, WIDTH, 1), device="cuda")
    tdx.mlp_init(P, tdx.LS_FEATURES, WIDTH, 1, seed)
    Xd = torch.from_numpy(np.ascontiguousarray(features, np.float32).ravel()).cuda()
    Td = torch.from_numpy(np.ascontiguousarray(labels, np.float32).ravel()).cuda()
    hist = tdx.mlp_train(P, Xd, Td, n, tdx.LS_FEATURES, WIDTH, 1, epochs, min(batch, n), lr, seed + 1)
    return P, hist


def sweeps_and_seconds(fit, ms, target):
    hit = np.nonzero(fit <= target)[0]
    if hit.size == 0:
        return None, None
    k = int(hit[0])
    return k + 1, float(np.sum(ms[:k + 1]) / 1e3)


def evaluate(seeds, params, dims=(96, 96, 96, 96), R=24, eta=1e-3, max_sweeps=400, log=None):
    """Per seed: sweeps and seconds to relerr <= eta + 1e-4 (the 0.1% noise floor + 1e-4)
    for plain ALS, the cpd_lsb rule and the learned step, from the same initial factors."""
    import torch
    target = eta + 1e-4
    X = torch.empty(int(np.prod(dims)), device="cuda")
    rows = []
    for i, sd in enumerate(seeds):
        swamp_tensor(sd, dims, R, eta=eta, out=X)
        row = {"seed": int(sd)}
        for name, mode in (("plain", tdx.LS_NONE), ("lsb", tdx.LS_LSB), ("learned", tdx.LS_LEARNED)):
            U = init_factors(sd, dims, R)
            fit, ms, tr = run(X, dims, R, U, mode, max_sweeps, params=params, target=target)
            k, sec = sweeps_and_seconds(fit, ms, target)
            row[name] = {"sweeps": k, "seconds": sec, "final_relerr": float(fit[-1]),
                         "accept_rate": float(np.mean(tr["accepted"])) if tr else None}
        rows.append(row)
        if log and (i + 1) % max(1, len(seeds) // 10) == 0:
            log(f"evaluate {i + 1}/{len(seeds)}")
    return rows, target


def summarize(rows):
    out = {}
    for name in ("plain", "lsb", "learned"):
        ks = [r[name]["sweeps"] for r in rows]
        ss = [r[name]["seconds"] for r in rows]
        reached = [k is not None for k in ks]
        out[name] = {"reached": int(sum(reached)), "n": len(rows),
                     "median_sweeps": float(np.median([k for k in ks if k is not None])) if any(reached) else None,
                     "median_seconds": float(np.median([s for s in ss if s is not None])) if any(reached) else None}
    both = [(r["plain"]["seconds"], r["learned"]["seconds"]) for r in rows
            if r["plain"]["seconds"] is not None and r["learned"]["seconds"] is not None]
    if both:
        out["median_speedup_learned_vs_plain"] = float(np.median([a / b for a, b in both]))
    return out


def save(params, path, meta):
    p = params.detach().cpu().numpy().astype(np.float32)
    p.tofile(path + ".bin")
    with open(path + ".json", "w") as f:
        json.dump(dict(meta, nin=tdx.LS_FEATURES, width=WIDTH, nout=1, count=int(p.size)), f, indent=1)


def load(path):
    import torch
    meta = json.load(open(path + ".json"))
    p = np.fromfile(path + ".bin", dtype=np.float32)
    assert p.size == meta["count"]
    return torch.from_numpy(p).cuda(), meta


def main():
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--train-seeds", type=int, default=2000)
    ap.add_argument("--eval-seeds", type=int, default=200)
    ap.add_argument("--dim", type=int, default=96)
    ap.add_argument("--rank", type=int, default=24)
    ap.add_argument("--sweeps", type=int, default=60)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--max-sweeps", type=int, default=400)
    ap.add_argument("--out", default="extrapolation_mlp")
    ap.add_argument("--eval-only", action="store_true",
                    help="evaluate the saved weights (--out) on the held-out seeds without training")
    args = ap.parse_args()
    dims = (args.dim,) * 4

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.time()
    if args.eval_only:
        P, meta = load(args.out)
        seeds = range(2000, 2000 + args.eval_seeds)
        rows, target = evaluate(seeds, P, dims, args.rank, max_sweeps=args.max_sweeps, log=log)
        print(json.dumps({"target_relerr": target, "weights": meta, "eval_seeds": [2000, 1999 + args.eval_seeds],
                          "summary": summarize(rows), "per_seed": rows, "wall_s": time.time() - t0}))
        return
    feats, labels = collect(range(args.train_seeds), dims, args.rank, args.sweeps, log)
    log(f"collected {feats.shape[0]} samples in {time.time() - t0:.1f}s; mean label {labels.mean():.3f}")
    P, hist = train(feats, labels, epochs=args.epochs)
    log(f"trained: loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    save(P, args.out, {"train_seeds": args.train_seeds, "dims": list(dims), "rank": args.rank,
                       "loss_first": float(hist[0]), "loss_last": float(hist[-1])})
    seeds = range(2000, 2000 + args.eval_seeds)
    rows, target = evaluate(seeds, P, dims, args.rank, max_sweeps=args.max_sweeps, log=log)
    print(json.dumps({"target_relerr": target, "summary": summarize(rows), "train_samples": int(feats.shape[0]),
                      "loss": [float(hist[0]), float(hist[-1])], "wall_s": time.time() - t0}))


if __name__ == "__main__":
    main()
