"""The extrapolation MLP's HIP forward, backward and Adam kernels against the FP64 CPU
oracle (gradients are FP32 on the device: relative tolerance 1e-4)."""
import numpy as np
import pytest

import oracle as O
import tdx
from helpers import dev, empty, host

pytestmark = pytest.mark.gpu

NIN, WIDTH, NOUT = 12, 64, 1


def params(rng):
    P = [rng.standard_normal((WIDTH, NIN)) * 0.3, rng.standard_normal(WIDTH) * 0.1,
         rng.standard_normal((WIDTH, WIDTH)) * 0.15, rng.standard_normal(WIDTH) * 0.1,
         rng.standard_normal((NOUT, WIDTH)) * 0.2, rng.standard_normal(NOUT) * 0.1]
    flat = np.concatenate([p.ravel() for p in P]).astype(np.float32)
    return P, flat


def split(flat):
    sizes = [WIDTH * NIN, WIDTH, WIDTH * WIDTH, WIDTH, NOUT * WIDTH, NOUT]
    out, o = [], 0
    for n in sizes:
        out.append(flat[o:o + n])
        o += n
    return out


def test_param_count():
    assert tdx.mlp_param_count(NIN, WIDTH, NOUT) == WIDTH * NIN + WIDTH + WIDTH * WIDTH + WIDTH + NOUT * WIDTH + NOUT
    assert tdx.mlp_param_count(0, WIDTH, NOUT) == -1


def test_forward_backward_match_oracle():
    rng = np.random.default_rng(0)
    P, flat = params(rng)
    Pd = [p.astype(np.float32).astype(np.float64) for p in P]
    batch = 97
    x = rng.standard_normal((batch, NIN)).astype(np.float32)
    t = rng.standard_normal((batch, NOUT)).astype(np.float32)
    y = empty(batch * NOUT)
    tdx.mlp_forward(dev(flat), dev(x.ravel()), batch, NIN, WIDTH, NOUT, y)
    yo = O.mlp_forward(x.astype(np.float64), Pd, NIN, WIDTH, NOUT)
    assert np.allclose(host(y).reshape(batch, NOUT), yo, rtol=1e-5, atol=1e-5)
    g = empty(flat.size)
    loss = empty(1)
    tdx.mlp_backward(dev(flat), dev(x.ravel()), dev(t.ravel()), batch, NIN, WIDTH, NOUT, g, loss)
    lo, go = O.mlp_backward(x.astype(np.float64), t.astype(np.float64), Pd, NIN, WIDTH, NOUT)
    assert np.isclose(host(loss)[0], lo, rtol=1e-5)
    for a, b in zip(split(host(g)), go):
        b = b.ravel()
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b) + 1e-7


def test_adam_matches_oracle():
    rng = np.random.default_rng(1)
    n = 1000
    p = rng.standard_normal(n)
    pd_, m, v = dev(p.astype(np.float32)), empty(n), empty(n)
    m.zero_()
    v.zero_()
    po, mo, vo = p.copy(), np.zeros(n), np.zeros(n)
    for step in range(1, 6):
        g = rng.standard_normal(n)
        tdx.mlp_adam(pd_, dev(g.astype(np.float32)), m, v, step, 1e-2)
        O.adam(po, g.astype(np.float32).astype(np.float64), mo, vo, step, 1e-2)
    assert np.allclose(host(pd_), po, rtol=1e-5, atol=1e-6)


def test_training_fits_a_smooth_target():
    rng = np.random.default_rng(2)
    ns = 4096
    x = rng.standard_normal((ns, NIN)).astype(np.float32)
    t = (np.tanh(x[:, :1] - 0.5 * x[:, 1:2]) + 0.1 * x[:, 2:3] ** 2).astype(np.float32)
    p = empty(tdx.mlp_param_count(NIN, WIDTH, NOUT))
    tdx.mlp_init(p, NIN, WIDTH, NOUT, 7)
    hist = tdx.mlp_train(p, dev(x.ravel()), dev(t.ravel()), ns, NIN, WIDTH, NOUT, 40, 256, 3e-3, 11)
    assert hist[-1] < 0.1 * hist[0]
    y = empty(ns)
    tdx.mlp_forward(p, dev(x.ravel()), ns, NIN, WIDTH, NOUT, y)
    assert np.mean((host(y) - t.ravel()) ** 2) < 0.02
