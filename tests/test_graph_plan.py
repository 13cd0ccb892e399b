"""Host compilation of graph products through the C ABI (no device work):
validation errors and engine limits map to the documented statuses."""
import pytest

import _graphs as G
from pricing_amd import abi, engine


def _compile_status(req):
    try:
        engine.PricingPlan.compile(req)
    except engine.EngineError as e:
        return e.status
    return abi.PE_OK


@pytest.fixture(scope="module")
def doc():
    return G.replay_request("pseudo_mc_full_risk")[1]


def test_valid_graphs_compile(doc):
    for g, o in (G.european(180, 100.0), G.asian([30, 60, 90], [1 / 3] * 3, 100.0),
                 G.barrier_out(90, [30, 60, 90], 100.0, 80.0, True)):
        assert _compile_status(G.graph_request(doc, g, o, 180)) == abi.PE_OK


def test_invalid_graphs_are_rejected(doc):
    g = G.G()
    g.nodes.append((0, "add", [1, 1], {}))
    g.nodes.append((1, "add", [0, 0], {}))
    assert _compile_status(G.graph_request(doc, g, 0, 10)) == abi.PE_ERR_INVALID_ARGUMENT      # cycle
    g = G.G()
    s = g.spot(90)
    g.nodes.append((s, "literal", [], {"value": 1.0}))
    assert _compile_status(G.graph_request(doc, g, s, 90)) == abi.PE_ERR_INVALID_ARGUMENT      # duplicate id
    g = G.G()
    s = g.spot(90)
    assert _compile_status(G.graph_request(doc, g, s, 60)) == abi.PE_ERR_INVALID_ARGUMENT      # paid before fixing
    g = G.G()
    s = g.spot(90)
    bad = g.push("log", (g.lit(-1.0),))
    assert _compile_status(G.graph_request(doc, g, g.push("add", (s, bad)), 90)) == abi.PE_ERR_INVALID_ARGUMENT
    g = G.G()
    s = g.spot(90)
    sm = g.push("smooth_indicator", (s,), half_width=0.0)
    assert _compile_status(G.graph_request(doc, g, sm, 90)) == abi.PE_ERR_INVALID_ARGUMENT


def test_live_value_limit(doc):
    # 40 observations all live until one final sum: more than 32 registers
    g = G.G()
    spots = [g.spot(10 * (i + 1)) for i in range(40)]
    acc = spots[-1]
    for s in reversed(spots[:-1]):
        acc = g.push("add", (s, acc))
    assert _compile_status(G.graph_request(doc, g, acc, 400)) == abi.PE_ERR_SIZE_LIMIT
    # the same sum accumulated in date order streams through two registers
    g = G.G()
    acc = g.spot(10)
    for i in range(1, 40):
        acc = g.push("add", (acc, g.spot(10 * (i + 1))))
    assert _compile_status(G.graph_request(doc, g, acc, 400)) == abi.PE_OK


def test_graph_greeks_compile_within_the_tangent_register_limit(doc):
    g, o = G.european(180, 100.0)
    assert _compile_status(G.graph_request(doc, g, o, 180, risk=True)) == abi.PE_OK
    # 24 observations live until one final sum: within the 32 registers of a
    # price, beyond the 20 that carry two tangents each for Greeks
    g = G.G()
    spots = [g.spot(10 * (i + 1)) for i in range(24)]
    acc = spots[-1]
    for s in reversed(spots[:-1]):
        acc = g.push("add", (s, acc))
    assert _compile_status(G.graph_request(doc, g, acc, 240)) == abi.PE_OK
    assert _compile_status(G.graph_request(doc, g, acc, 240, risk=True)) == abi.PE_ERR_UNSUPPORTED


def test_graph_greeks_under_local_vol_are_unsupported():
    _, doc = G.replay_request("pseudo_mc_price_only", "local_volatility")
    g, o = G.european(365, 100.0)
    req = G.graph_request(doc, g, o, 365)
    req.c.risk.delta = 1
    assert _compile_status(req) == abi.PE_ERR_UNSUPPORTED
