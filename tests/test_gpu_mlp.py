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


# GELU(x) = x Phi(x) and GELU(GELU(x)) at 40 digits (mpmath), the same known answers the
# oracle is pinned by in tests/test_oracle.py
GELU_KAT = [(0.0, 0.0), (1.0, 0.84134474606854294859), (-1.0, -0.15865525393145705141),
            (2.0, 1.9544997361036415856), (-3.0, -0.00404969409489028358), (0.5, 0.34573123063700655182)]
GELU_GELU_KAT = [(1.0, 0.67301066416940404567), (-1.0, -0.069327625476641394937), (2.0, 1.9050097055649646399)]


def _packed(W1, b1, W2, b2, W3, b3, nout):
    assert W1.shape == (WIDTH, NIN) and W3.shape == (nout, WIDTH)
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)


def test_forward_gelu_known_answers():
    """Selector weights isolate single GELUs on the device: W2 = 0, b2[k] = x_k and W3 = the
    first rows of the identity give y_k = GELU(x_k); unit diagonal weights give GELU(GELU(x_k))."""
    nout = len(GELU_KAT)
    sel = np.zeros((nout, WIDTH))
    sel[np.arange(nout), np.arange(nout)] = 1.0
    b2 = np.zeros(WIDTH)
    b2[:nout] = [x for x, _ in GELU_KAT]
    flat = _packed(np.zeros((WIDTH, NIN)), np.zeros(WIDTH), np.zeros((WIDTH, WIDTH)), b2, sel, np.zeros(nout), nout)
    y = empty(nout)
    tdx.mlp_forward(dev(flat), dev(np.zeros(NIN, np.float32)), 1, NIN, WIDTH, nout, y)
    assert np.allclose(host(y), [v for _, v in GELU_KAT], rtol=1e-6, atol=1e-7), host(y)
    n2 = len(GELU_GELU_KAT)
    W1 = np.zeros((WIDTH, NIN))
    W1[np.arange(n2), np.arange(n2)] = 1.0
    W2 = np.eye(WIDTH)
    flat = _packed(W1, np.zeros(WIDTH), W2, np.zeros(WIDTH), sel[:n2], np.zeros(n2), n2)
    x = np.zeros(NIN, np.float32)
    x[:n2] = [v for v, _ in GELU_GELU_KAT]
    y = empty(n2)
    tdx.mlp_forward(dev(flat), dev(x), 1, NIN, WIDTH, n2, y)
    assert np.allclose(host(y), [v for _, v in GELU_GELU_KAT], rtol=1e-6, atol=1e-7), host(y)


def test_adam_two_step_known_answer():
    """Kingma & Ba (2015): lr 0.1, p0 = 1, gradients 0.5 then -0.25 -> 0.900000002, then
    0.87336629870784616 (40-digit mpmath values; FP32 on the device)."""
    p, m, v = dev(np.ones(4, np.float32)), empty(4), empty(4)
    m.zero_()
    v.zero_()
    tdx.mlp_adam(p, dev(np.full(4, 0.5, np.float32)), m, v, 1, 0.1)
    assert np.allclose(host(p), 0.90000000199999996, rtol=0, atol=1e-7)
    tdx.mlp_adam(p, dev(np.full(4, -0.25, np.float32)), m, v, 2, 0.1)
    assert np.allclose(host(p), 0.87336629870784616256, rtol=0, atol=2e-7)
    assert np.allclose(host(m), 0.02, rtol=1e-6) and np.allclose(host(v), 3.1225e-4, rtol=1e-6)
