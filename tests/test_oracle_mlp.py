"""Pins the oracle's continuation network (2 -> 64 -> 64 -> 1, SiLU, MSE):
analytic gradients against central finite differences of the loss, the
forward pass against a direct numpy evaluation, and the bfloat16 rounding
against its definition (round to nearest even on the upper 16 bits).
"""
import ctypes as C

import numpy as np
import pytest

from _support import oracle

P = 4417


def _params(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, P)
    p[128:192] = rng.uniform(-0.1, 0.1, 64)
    return p


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(C.POINTER(C.c_double)), a


def _grad(p, x, y, bf16=0):
    g = np.zeros(P)
    pp, _ = _d(p)
    xp, _x = _d(x)
    yp, _y = _d(y)
    loss = oracle().oracle_mlp_gradients(pp, xp, yp, len(y), bf16, g.ctypes.data_as(C.POINTER(C.c_double)))
    return loss, g


def _numpy_forward(p, x):
    w1 = p[0:128].reshape(64, 2)
    b1 = p[128:192]
    w2 = p[192:4288].reshape(64, 64)
    b2 = p[4288:4352]
    w3 = p[4352:4416]
    b3 = p[4416]
    silu = lambda a: a / (1.0 + np.exp(-a))
    h1 = silu(x @ w1.T + b1)
    h2 = silu(h1 @ w2.T + b2)
    return h2 @ w3 + b3


def test_forward_matches_numpy():
    p = _params(1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(50, 2))
    out = np.zeros(50)
    pp, _ = _d(p)
    xp, _x = _d(x)
    oracle().oracle_mlp_forward(pp, xp, 50, 0, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.allclose(out, _numpy_forward(p, x), rtol=1e-13, atol=1e-13)


def test_gradient_matches_finite_differences():
    p = _params(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(12, 2))
    y = rng.normal(size=12)
    loss, g = _grad(p, x, y)
    assert loss == pytest.approx(np.mean((_numpy_forward(p, x) - y) ** 2), rel=1e-13)
    idx = list(range(0, 128, 17)) + list(range(128, 192, 13)) + list(range(192, 4288, 331)) + \
        list(range(4288, 4352, 11)) + list(range(4352, 4416, 9)) + [4416]
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pu, pd = p.copy(), p.copy()
        pu[i] += h
        pd[i] -= h
        fd = (np.mean((_numpy_forward(pu, x) - y) ** 2) - np.mean((_numpy_forward(pd, x) - y) ** 2)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)), i


def test_bf16_rounding():
    import struct
    lib = oracle()
    p = np.zeros(P)
    # W2 entries that are not bfloat16 numbers: the bf16 forward differs and
    # equals the forward with W2 rounded by hand (nearest even on 16 bits)
    rng = np.random.default_rng(5)
    p[:] = _params(6)
    x = rng.normal(size=(8, 2))

    def rnd(v):
        u = struct.unpack("<I", struct.pack("<f", v))[0]
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return struct.unpack("<f", struct.pack("<I", u))[0]

    assert rnd(1.0 + 2 ** -8) == 1.0 and rnd(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6 and rnd(1.0 + 2 ** -7) == 1.0 + 2 ** -7
    out = np.zeros(8)
    pp, _ = _d(p)
    xp, _x = _d(x)
    lib.oracle_mlp_forward(pp, xp, 8, 1, out.ctypes.data_as(C.POINTER(C.c_double)))
    exact = _numpy_forward(p, x)
    assert np.max(np.abs(out - exact)) > 1e-6

    def rnd_arr(a):
        u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
        return u.view(np.float32).astype(np.float64)

    w1 = p[0:128].reshape(64, 2)
    silu = lambda a: a / (1.0 + np.exp(-a))
    h1 = rnd_arr(silu(x @ w1.T + p[128:192]))
    w2 = rnd_arr(p[192:4288]).reshape(64, 64)
    want = silu(h1 @ w2.T + p[4288:4352]) @ p[4352:4416] + p[4416]
    assert np.allclose(out, want, rtol=1e-12, atol=1e-12)
