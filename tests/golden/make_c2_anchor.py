"""Writes tests/golden/c2_fp64_anchor.json: the FP64 oracle's CP-ALS trajectory on the
1024 x 1024 x 1024 rank-64 tensor (N(0,1) factors with unit columns from Philox seed 0, 0.1%
Gaussian noise from seed 2), 27 sweeps from the N(0,1) initial factors of each seed.  These
are the accuracy anchors of the mixed-precision benchmark: a path reaches its target when its
relative error is within 1e-4 of the FP64 oracle's after 27 sweeps from the same init.

The tensor and the initial factors are generated on the device by libtdx (the benchmark's
own recipe) and copied to the host, so the oracle runs on exactly the FP32 values the GPU
paths see.  The ALS itself is oracle.cp_als3_gemm in FP64 (the GEMM formulation of the
reference CPU path, decomposition/CPU/python/mycp.py:11-22), which tests/test_oracle.py pins
against orc_cp_als.  Needs a GPU (for the generator) and ~13 GiB of host memory; about 1-3 s
per FP64 sweep on 8-16 host cores.

Run:  python tests/golden/make_c2_anchor.py --out gpurun_out/c2_fp64_anchor.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "first-project_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def fingerprint(Xh):
    """||X||^2 in FP64 and 8 sampled entries: identifies the generated tensor."""
    idx = np.linspace(0, Xh.size - 1, 8).astype(np.int64)
    return {"normx2": float(np.dot(Xh.astype(np.float64), Xh.astype(np.float64))),
            "sample_index": [int(i) for i in idx], "sample": [float(Xh[i]) for i in idx]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", default="1,2,3,4,5,6,7,8,9")
    ap.add_argument("--sweeps", type=int, default=27)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--rank", type=int, default=64)
    ap.add_argument("--out", default=os.path.join(HERE, "c2_fp64_anchor.json"))
    a = ap.parse_args()

    import torch

    import oracle as O
    import tdx
    import workloads as W

    def log(m):
        print(m, flush=True)

    tdx.lib()
    X, dims, R = W.c2(a.size, a.rank)
    Xh = X.cpu().numpy()
    del X
    torch.cuda.empty_cache()
    out = {"workload": f"CP-ALS {a.size}^3 rank {R}, N(0,1) unit-column factors (seed 0), 0.1% noise (seed 2), "
                       f"N(0,1) init per seed; FP64 oracle (GEMM formulation), {a.sweeps} sweeps",
           "dims": list(dims), "rank": R, "sweeps": a.sweeps, "tensor": fingerprint(Xh),
           "threads": os.environ.get("OMP_NUM_THREADS"), "seeds": {}}
    log(f"tensor ready, ||X||^2 = {out['tensor']['normx2']:.9e}")
    X64 = Xh.astype(np.float64)
    del Xh
    for seed in (int(s) for s in a.seeds.split(",")):
        U0 = [u.cpu().numpy().reshape((d, R), order="F").astype(np.float64)
              for u, d in zip(W.normal_init(dims, R, seed), dims)]
        t0 = time.perf_counter()
        U, fit, secs = O.cp_als3_gemm(X64, dims, U0, a.sweeps, log=log)
        log(f"seed {seed}: relerr after {a.sweeps} sweeps {fit[-1]:.12f} ({time.perf_counter() - t0:.1f} s); "
            f"trajectory {np.round(fit, 6).tolist()}")
        out["seeds"][str(seed)] = {"relerr": [float(v) for v in fit], "seconds": [float(v) for v in secs]}
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    log(f"wrote {a.out}")


if __name__ == "__main__":
    main()
