"""The device continuation network (BF16 matrix-core forward/backward, FP32
accumulation) against the FP64 oracle network.

Two bars: against the oracle with the device's bfloat16 rounding emulated
(W2, hidden activations, layer-2 gradients) the device agrees to FP32
accumulation error (relative L2 1e-3 per parameter block); against the exact
FP64 network it agrees to bfloat16 precision (relative L2 3e-2), and the
cosine of the full gradient vectors exceeds 0.999.
"""
import ctypes as C

import numpy as np
import pytest

from _support import oracle
from pricing_amd import engine

pytestmark = pytest.mark.gpu

P = 4417
BLOCKS = {"W1": (0, 128), "b1": (128, 192), "W2": (192, 4288), "b2": (4288, 4352), "w3": (4352, 4416),
          "b3": (4416, 4417)}


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(P)
    p[0:128] = rng.uniform(-1, 1, 128) * np.sqrt(6.0 / 2)
    p[128:192] = rng.uniform(-0.1, 0.1, 64)
    p[192:4288] = rng.uniform(-1, 1, 4096) * np.sqrt(6.0 / 64)
    p[4288:4352] = rng.uniform(-0.1, 0.1, 64)
    p[4352:4416] = rng.uniform(-1, 1, 64) * np.sqrt(3.0 / 64)
    p[4416] = 0.05
    return p.astype(np.float32).astype(np.float64)


def _fp(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def device_gradient(p, x, y):
    pk, pp = _fp(p)
    xk, xp = _fp(x)
    yk, yp = _fp(y)
    g = np.zeros(P, dtype=np.float32)
    loss = C.c_double()
    engine._check(engine.load().pe_mlp_gradients(pp, xp, yp, len(y), g.ctypes.data_as(C.POINTER(C.c_float)),
                                                  C.byref(loss)))
    return loss.value, g.astype(np.float64)


def oracle_gradient(p, x, y, bf16):
    pk, pp = _dp(p)
    xk, xp = _dp(np.asarray(x, dtype=np.float32))
    yk, yp = _dp(np.asarray(y, dtype=np.float32))
    g = np.zeros(P)
    loss = oracle().oracle_mlp_gradients(pp, xp, yp, len(y), int(bf16), g.ctypes.data_as(C.POINTER(C.c_double)))
    return loss, g


@pytest.mark.parametrize("rows", [8192, 1000, 37])
def test_gradient_matches_oracle(rows):
    rng = np.random.default_rng(rows)
    p = _params(rows + 1)
    x = rng.normal(size=(rows, 2))
    y = rng.normal(size=rows)
    dl, dg = device_gradient(p, x, y)
    el, eg = oracle_gradient(p, x, y, bf16=True)
    fl, fg = oracle_gradient(p, x, y, bf16=False)
    assert abs(dl - el) <= 1e-4 * abs(el)
    for name, (a, b) in BLOCKS.items():
        ref = eg[a:b]
        err = np.linalg.norm(dg[a:b] - ref) / max(np.linalg.norm(ref), 1e-30)
        assert err < 1e-3, (name, err)
        exact = np.linalg.norm(dg[a:b] - fg[a:b]) / max(np.linalg.norm(fg[a:b]), 1e-30)
        assert exact < 3e-2, (name, exact)
    assert np.dot(dg, fg) / (np.linalg.norm(dg) * np.linalg.norm(fg)) > 0.999


def test_forward_matches_oracle():
    rng = np.random.default_rng(9)
    p = _params(10)
    x = rng.normal(size=(333, 2))
    pk, pp = _fp(p)
    xk, xp = _fp(x)
    out = np.zeros(333, dtype=np.float32)
    engine._check(engine.load().pe_mlp_forward(pp, xp, 333, out.ctypes.data_as(C.POINTER(C.c_float))))
    ref = np.zeros(333)
    pd, pdp = _dp(p)
    xd, xdp = _dp(np.asarray(x, dtype=np.float32))
    oracle().oracle_mlp_forward(pdp, xdp, 333, 1, ref.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.max(np.abs(out - ref)) <= 1e-4 * max(1.0, np.max(np.abs(ref)))
