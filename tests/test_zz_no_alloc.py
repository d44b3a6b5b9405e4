"""tdx_mttkrp_ex_noalloc, the MTTKRP entry point that allocates nothing (no counterpart in the
reference; for callers that cannot allocate device memory), against tdx_mttkrp_ex: the same
kernels and launch plan, so the outputs must be equal bit for bit.  Two calls in a row, the
second smaller (the fixed scratch keeps the first call's contents), for FP32, BF16 and FP8
storage; an input whose panels and partials exceed the 64 MiB scratch returns
TDX_ERR_TOO_LARGE before any device work; and the call leaves the device's free memory as it
was, where tdx_mttkrp_ex grows its workspace pool."""
import numpy as np
import pytest

import oracle as O
import tdx
from helpers import dev, empty, host

pytestmark = pytest.mark.gpu


def _storage(X, dtype, I0):
    import torch
    if dtype == tdx.TDX_F32:
        return X, None
    if dtype == tdx.TDX_BF16:
        Xb = torch.empty(X.numel(), dtype=torch.int16, device="cuda")
        tdx.quantize_bf16(X, Xb)
        return Xb, None
    Xq = torch.empty(X.numel(), dtype=torch.uint8, device="cuda")
    sc = empty(I0)
    tdx.quantize_fp8_mode1(X, Xq, sc, I0)
    return Xq, sc


@pytest.mark.parametrize("dtype", [tdx.TDX_F32, tdx.TDX_BF16, tdx.TDX_FP8_E4M3], ids=["fp32", "bf16", "fp8"])
def test_noalloc_matches_mttkrp_ex_on_two_calls(dtype):
    rng = np.random.default_rng(3)
    for dims, R, mode in (((256, 192, 160), 48, 1), ((64, 48, 40), 16, 0)):
        print("noalloc", dtype, dims, R, mode, flush=True)
        X = dev(rng.standard_normal(int(np.prod(dims))).astype(np.float32))
        Xs, sc = _storage(X, dtype, dims[0])
        F = [dev(rng.standard_normal(d * R).astype(np.float32)) for d in dims]
        ref, out = empty(dims[mode] * R), empty(dims[mode] * R)
        out.fill_(float("nan"))
        tdx.mttkrp_ex(Xs, dtype, sc, dims, F, R, mode, ref)
        st = tdx.mttkrp_ex_noalloc(Xs, dtype, sc, dims, F, R, mode, out)
        assert st == tdx.TDX_OK
        assert np.array_equal(host(out), host(ref))
        if dtype == tdx.TDX_F32:
            o = O.mttkrp(host(X), dims, [host(f).reshape((d, R), order="F") for f, d in zip(F, dims)], mode)
            assert np.linalg.norm(host(out).reshape(o.shape, order="F") - o) <= 1e-5 * np.linalg.norm(o)


def test_noalloc_rejects_oversize_before_device_work():
    """A 2^20-long mode makes one factor panel 64 MiB at rank 16: TDX_ERR_TOO_LARGE, and the
    output (NaN-filled) is not written."""
    X = empty(64)
    F = [empty(64) for _ in range(3)]
    out = empty(16 * 16)
    out.fill_(float("nan"))
    st = tdx.mttkrp_ex_noalloc(X, tdx.TDX_F32, None, (16, 1 << 20, 4), F, 16, 0, out)
    assert st == tdx.TDX_ERR_TOO_LARGE
    assert tdx.lib().tdx_last_status() == tdx.TDX_ERR_TOO_LARGE
    assert "no-allocation" in tdx.lib().tdx_status_string(st).decode()
    assert np.all(np.isnan(host(out)))


def test_noalloc_does_not_grow_device_memory():
    import torch
    dims, R, mode = (512, 384, 256), 64, 0
    rng = np.random.default_rng(4)
    X = dev(rng.standard_normal(int(np.prod(dims))).astype(np.float32))
    F = [dev(rng.standard_normal(d * R).astype(np.float32)) for d in dims]
    out = empty(dims[mode] * R)
    assert tdx.mttkrp_ex_noalloc(X, tdx.TDX_F32, None, dims, F, R, mode, out) == tdx.TDX_OK  # warm-up
    torch.cuda.synchronize()
    tdx.lib().tdx_release_workspace()
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        assert tdx.mttkrp_ex_noalloc(X, tdx.TDX_F32, None, dims, F, R, mode, out) == tdx.TDX_OK
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    tdx.mttkrp_ex(X, tdx.TDX_F32, None, dims, F, R, mode, out)
    torch.cuda.synchronize()
    free2, _ = torch.cuda.mem_get_info()
    print(f"free memory: before {free0}, after the no-allocation calls {free1}, after tdx_mttkrp_ex {free2}")
    assert free0 - free1 < (1 << 20)
    assert free1 - free2 >= (4 << 20)  # the pooled entry point does allocate: the measure sees it
