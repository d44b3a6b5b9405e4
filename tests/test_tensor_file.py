"""The quantized tensor file format (.tdxq; the step before MTTKRP, extending the reference's
in-memory precision reduction f2h, src/float2half.cu:14-18).  CPU tests write and load host
buffers through the C ABI and read the bytes back independently against the documented layout
(include/tdx.h); the GPU test stages device buffers and runs the MTTKRP on the loaded tensor."""
import struct

import numpy as np
import pytest

import oracle as O
import tdx


def _fnv1a(b):
    h = 1469598103934665603
    for x in b:
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _host_tensor(storage, dims, seed=0):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal(int(np.prod(dims))) * 3).astype(np.float32)
    if storage == tdx.TDX_F32:
        return X, X, None
    if storage == tdx.TDX_BF16:
        return X, O.bf16_bits(X), None
    Q, s = O.quantize_e4m3_mode1(X, dims[0])
    return X, Q, s


@pytest.mark.parametrize("storage", [tdx.TDX_F32, tdx.TDX_BF16, tdx.TDX_FP8_E4M3], ids=["fp32", "bf16", "fp8"])
@pytest.mark.parametrize("dims", [(7, 5, 3), (12, 6, 5, 4)], ids=["order3", "order4"])
def test_round_trip_and_layout(tmp_path, storage, dims):
    _, data, scales = _host_tensor(storage, dims)
    path = str(tmp_path / "x.tdxq")
    tdx.tensor_save(path, data, storage, dims, scales)
    raw = open(path, "rb").read()
    # the documented layout, read independently
    assert raw[:4] == b"TDXQ" and struct.unpack_from("<III", raw, 4) == (1, len(dims), storage)
    fdims = struct.unpack_from("<4q", raw, 16)
    assert fdims == tuple(list(dims) + [1] * (4 - len(dims)))
    ns, nbytes, chk = struct.unpack_from("<qqQ", raw, 48)
    assert chk == _fnv1a(raw[:64]) and raw[72:128] == bytes(56)
    esz = {tdx.TDX_F32: 4, tdx.TDX_BF16: 2, tdx.TDX_FP8_E4M3: 1}[storage]
    assert nbytes == int(np.prod(dims)) * esz
    off = 128
    if storage == tdx.TDX_FP8_E4M3:
        assert ns == dims[0]
        assert np.array_equal(np.frombuffer(raw, np.float32, ns, off), scales)
        off += -(-4 * ns // 128) * 128
    else:
        assert ns == 0
    assert raw[off:] == data.tobytes() and len(raw) == off + nbytes
    # loader: header, data and scales back bit for bit
    info = tdx.tensor_info(path)
    assert (info.order, info.storage, info.n_scales, info.data_bytes) == (len(dims), storage, ns, nbytes)
    back = np.zeros_like(data)
    sb = np.zeros(dims[0], np.float32) if scales is not None else None
    tdx.tensor_load(path, back, sb)
    assert back.tobytes() == data.tobytes()
    if scales is not None:
        assert np.array_equal(sb, scales)


def test_corrupt_and_short_files_are_rejected(tmp_path):
    dims = (8, 4, 2)
    _, data, scales = _host_tensor(tdx.TDX_FP8_E4M3, dims)
    path = str(tmp_path / "x.tdxq")
    tdx.tensor_save(path, data, tdx.TDX_FP8_E4M3, dims, scales)
    raw = bytearray(open(path, "rb").read())
    for pos in (0, 9, 20, 50, 60):  # magic, order, a dim, n_scales, data_bytes: the checksum catches each
        bad = bytearray(raw)
        bad[pos] ^= 0x01
        p = str(tmp_path / f"bad{pos}.tdxq")
        open(p, "wb").write(bytes(bad))
        with pytest.raises(tdx.TdxError) as e:
            tdx.tensor_info(p)
        assert e.value.status == tdx.TDX_ERR_IO
    short = str(tmp_path / "short.tdxq")
    open(short, "wb").write(bytes(raw[:-3]))
    with pytest.raises(tdx.TdxError) as e:
        tdx.tensor_load(short, np.zeros_like(data), np.zeros(8, np.float32))
    assert e.value.status == tdx.TDX_ERR_IO
    with pytest.raises(tdx.TdxError) as e:  # destination too small
        tdx.tensor_load(path, np.zeros(data.size - 1, np.uint8), np.zeros(8, np.float32))
    assert e.value.status == tdx.TDX_ERR_ARG
    with pytest.raises(tdx.TdxError) as e:  # missing scales for FP8
        tdx.tensor_load(path, np.zeros_like(data), None)
    assert e.value.status == tdx.TDX_ERR_ARG
    with pytest.raises(tdx.TdxError) as e:
        tdx.tensor_info(str(tmp_path / "missing.tdxq"))
    assert e.value.status == tdx.TDX_ERR_IO


@pytest.mark.gpu
@pytest.mark.parametrize("storage", [tdx.TDX_BF16, tdx.TDX_FP8_E4M3], ids=["bf16", "fp8"])
def test_device_save_load_then_mttkrp(tmp_path, storage):
    """Quantize on the device, save from device memory, load into fresh device buffers: the bytes
    and the scales come back bit for bit and the MTTKRP of the loaded tensor equals the MTTKRP
    of the in-memory quantized tensor bit for bit.  The tensor (96 x 80 x 72, 2.1 MiB in BF16)
    spans several staging chunks only in FP32, so a second FP32 save/load of 24 M elements
    (96 MiB, two 64 MiB chunks) checks the chunked path."""
    import torch

    from helpers import dev, empty, host
    dims, R = (96, 80, 72), 16
    rng = np.random.default_rng(1)
    X = dev(rng.standard_normal(int(np.prod(dims))).astype(np.float32))
    F = [dev(rng.standard_normal(d * R).astype(np.float32)) for d in dims]
    n = X.numel()
    if storage == tdx.TDX_BF16:
        Xq = torch.empty(n, dtype=torch.int16, device="cuda")
        tdx.quantize_bf16(X, Xq)
        sc = None
    else:
        Xq = torch.empty(n, dtype=torch.uint8, device="cuda")
        sc = empty(dims[0])
        tdx.quantize_fp8_mode1(X, Xq, sc, dims[0])
    path = str(tmp_path / "q.tdxq")
    tdx.tensor_save(path, Xq, storage, dims, sc)
    Xl = torch.empty_like(Xq)
    scl = empty(dims[0]) if sc is not None else None
    info = tdx.tensor_load(path, Xl, scl)
    assert info.storage == storage and tuple(info.dims[:3]) == dims
    assert torch.equal(Xl, Xq)
    if sc is not None:
        assert torch.equal(scl, sc)
    for mode in range(3):
        a, b = empty(dims[mode] * R), empty(dims[mode] * R)
        tdx.mttkrp_ex(Xq, storage, sc, dims, F, R, mode, a)
        tdx.mttkrp_ex(Xl, storage, scl, dims, F, R, mode, b)
        assert torch.equal(a, b)
    big = dev(rng.standard_normal(24 << 20).astype(np.float32))
    p2 = str(tmp_path / "big.tdxq")
    tdx.tensor_save(p2, big, tdx.TDX_F32, (1024, 4096, 6))
    back = torch.empty_like(big)
    tdx.tensor_load(p2, back)
    assert torch.equal(back, big)
    del host
