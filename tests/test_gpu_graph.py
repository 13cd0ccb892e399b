"""GPU parity of graph products (the register tape in the BS batch kernel
and in the graph flavour of the local-vol kernel) against the oracle's lazy
graph evaluation: FP64 device mode within 1e-12 of the oracle (and of the
replay fixtures' prices for the graph forms of their products, Black-Scholes
and local vol), the mixed mode within 0.1 standard errors; the device's
poison bits raise exactly where the reference's lazy evaluation fails."""
import copy

import pytest

import _graphs as G
from _support import oracle_evaluate
from pricing_amd import abi, engine, request

pytestmark = pytest.mark.gpu

REL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def _close(a, b, rel=REL):
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-300)


def _check(req, block=0, mixed=True):
    ref = oracle_evaluate(req, workers=4, block=block)
    plan = engine.PricingPlan.compile(req, request.exec_policy(1, block) if block else None)
    dev = plan.evaluate("fp64")
    print("graph", ref.estimates[0].value, dev.estimates[0].value, flush=True)
    assert _close(dev.estimates[0].value, ref.estimates[0].value), (dev.estimates[0].value, ref.estimates[0].value)
    assert _close(dev.estimates[0].standard_error, ref.estimates[0].standard_error, 1e-10)
    if mixed:
        mx = plan.evaluate("mixed")
        se = ref.estimates[0].standard_error
        assert abs(mx.estimates[0].value - ref.estimates[0].value) <= 0.1 * se + 1e-12
    return dev


@pytest.mark.parametrize("name", ["pseudo_mc_full_risk", "rqmc_full_risk", "digital_exact_price_only"])
def test_fixture_products_as_graphs(name):
    from test_oracle_graph import fixture_graph

    case, doc = G.replay_request(name)
    g, out, pay = fixture_graph(doc)
    dev = _check(G.graph_request(doc, g, out, pay), block=case["reduction_block_size"], mixed=False)
    assert _close(dev.estimates[0].value, case["value"]["value"])


@pytest.mark.parametrize("engine_kind", ["pseudo", "rqmc"])
def test_path_dependent_graphs(engine_kind):
    _, doc = G.replay_request("pseudo_mc_full_risk" if engine_kind == "pseudo" else "rqmc_full_risk")
    days = [30 * (i + 1) for i in range(12)]
    for g, o in (G.asian(days, [1.0 / 12] * 12, 100.0), G.lookback(days, 95.0, call=False),
                 G.barrier_out(days[-1], days, 100.0, 85.0, down=True)):
        _check(G.graph_request(doc, g, o, days[-1]))


@pytest.mark.parametrize("name", ["digital.json", "coupon_and_capped_call.golden.json"])
def test_product_graph_documents(name):
    _, doc = G.replay_request("pseudo_mc_full_risk")
    req = G.graph_request(doc, G.G(), 0, 0)
    req.set_product_graph_document(G.golden_graph(name), G.dt.date.fromisoformat(doc["valuation_date"]))
    _check(req)


def test_generated_dags_with_selects():
    _, doc = G.replay_request("pseudo_mc_full_risk")
    days = (60, 120, 180)
    ran = 0
    for ref, outputs in G.lcg_dags(count=12, ops=20):
        g, o = G.dag_nodes(ref, outputs[0], days)
        # every observation feeds the output through the tape, one cash flow
        _check(G.graph_request(doc, g, o, days[-1]), mixed=False)
        ran += 1
    assert ran == 12


def test_untaken_log_is_no_error_and_a_live_one_is():
    _, doc = G.replay_request("pseudo_mc_full_risk")
    g = G.G()
    s = g.spot(180)
    k = g.lit(150.0)
    c = g.push("compare", (s, k), operation="greater")
    lg = g.push("log", (g.push("subtract", (s, k)),))
    out = g.push("select", (c, lg, g.lit(0.0)))
    _check(G.graph_request(doc, g, out, 180))
    g2 = G.G()
    s2 = g2.spot(180)
    live = g2.push("maximum", (g2.push("log", (g2.push("subtract", (s2, g2.lit(150.0))),)), g2.lit(0.0)))
    plan = engine.PricingPlan.compile(G.graph_request(doc, g2, live, 180))
    with pytest.raises(engine.EngineError) as e:
        plan.evaluate("fp64")
    assert e.value.status == abi.PE_ERR_NONFINITE_STATE
    with pytest.raises(RuntimeError):
        oracle_evaluate(G.graph_request(doc, g2, live, 180))


def test_pre_dividend_spot_on_an_ex_date():
    _, doc = G.replay_request("pseudo_mc_full_risk")
    doc = copy.deepcopy(doc)
    doc["market"]["discrete_dividends"] = [{"ex_time": 120 / 365, "quote": {"type": "fixed_cash", "amount": 2.0}}]
    g = G.G()
    pre = g.pre_spot(120)
    post = g.spot(120)
    term = g.spot(240)
    drop = g.push("subtract", (pre, post))
    call = g.push("maximum", (g.push("subtract", (term, g.lit(100.0))), g.lit(0.0)))
    out = g.push("add", (call, drop))
    _check(G.graph_request(doc, g, out, 240))


# ---- graph products on local-vol paths (the graph flavour of lv_kernel) ----

@pytest.mark.parametrize("name", ["pseudo_mc_price_only", "rqmc_price_only"])
def test_local_vol_fixture_products_as_graphs(name):
    """The local-vol replay fixtures' European as a graph: the FP64 device
    mode reproduces the oracle and the fixture's recorded price."""
    from test_oracle_graph import fixture_graph

    case, doc = G.replay_request(name, "local_volatility")
    g, out, pay = fixture_graph(doc)
    dev = _check(G.graph_request(doc, g, out, pay), block=case["reduction_block_size"])
    assert _close(dev.estimates[0].value, case["value"]["value"])


def test_local_vol_path_dependent_graphs():
    """Monthly Asian, lookback and barrier graphs on local-vol paths: FP64
    device = oracle within 1e-12, mixed within 0.1 SE; the graph Asian equals
    the fixed Asian on the same device paths in FP64 mode."""
    _, doc = G.replay_request("pseudo_mc_price_only", "local_volatility")
    days = [30 * (i + 1) for i in range(12)] + [365]
    w = [1.0 / 13] * 13
    for g, o in (G.asian(days, w, 100.0), G.lookback(days, 95.0, call=False),
                 G.barrier_out(days[-1], days, 100.0, 85.0, down=True)):
        _check(G.graph_request(doc, g, o, days[-1]))
    g, o = G.asian(days, w, 100.0)
    graph = engine.PricingPlan.compile(G.graph_request(doc, g, o, days[-1])).evaluate("fp64")
    plain = request.Request.from_v3(copy.deepcopy(doc))
    plain.c.risk.delta = plain.c.risk.vega = plain.c.risk.gamma = 0
    plain.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, days[-1], 100.0, dates=days, weights=w)
    fixed = engine.PricingPlan.compile(plain).evaluate("fp64")
    assert _close(graph.estimates[0].value, fixed.estimates[0].value)


def test_local_vol_graph_non_finite_value_fails():
    """A graph whose evaluated node is non-finite (log of a negative
    difference) fails the unit under local vol in both modes."""
    _, doc = G.replay_request("pseudo_mc_price_only", "local_volatility")
    g = G.G()
    s = g.spot(365)
    bad = g.push("log", (g.push("subtract", (s, g.lit(1e9))),))
    plan = engine.PricingPlan.compile(G.graph_request(doc, g, bad, 365))
    for mode in ("fp64", "mixed"):
        with pytest.raises(engine.EngineError) as err:
            plan.evaluate(mode)
        assert err.value.status == abi.PE_ERR_NONFINITE_STATE


# ---- pathwise Greeks of graph products (forward tangents in LDS) ----

def _check_risk(req, block=0, mixed=True):
    """All ten risk components (price, pathwise Delta / Vega, Gamma from the
    bumped Deltas, the bump validations) of the FP64 device mode against the
    oracle's reverse sweep, 1e-12 of each component's conditioning (a
    validation component divides price differences by the bump); the mixed
    mode's price, Delta and Vega within 0.1 standard errors."""
    from test_gpu_parity import compare

    ref = oracle_evaluate(req, workers=4, block=block)
    plan = engine.PricingPlan.compile(req, request.exec_policy(1, block) if block else None)
    dev = plan.evaluate("fp64")
    for c in range(10):
        print("graph risk", abi.COMPONENT_NAMES[c], ref.estimates[c].value, dev.estimates[c].value, flush=True)
    compare(dev, ref, list(range(10)), req=req)
    if mixed:
        mx = plan.evaluate("mixed")
        for c in (0, 1, 2):
            se = ref.estimates[c].standard_error
            assert abs(mx.estimates[c].value - ref.estimates[c].value) <= 0.1 * se + 1e-12, abi.COMPONENT_NAMES[c]
    return dev


@pytest.mark.parametrize("name", ["pseudo_mc_full_risk", "rqmc_full_risk"])
def test_fixture_greeks_as_graphs(name):
    from test_oracle_graph import fixture_graph

    case, doc = G.replay_request(name)
    g, out, pay = fixture_graph(doc)
    dev = _check_risk(G.graph_request(doc, g, out, pay, risk=True), block=case["reduction_block_size"], mixed=False)
    for key, index in (("delta", 1), ("vega", 2), ("gamma", 3)):
        assert _close(dev.estimates[index].value, case["risks"][key]["value"]), key


@pytest.mark.parametrize("engine_kind", ["pseudo", "rqmc"])
def test_path_dependent_graph_greeks(engine_kind):
    _, doc = G.replay_request("pseudo_mc_full_risk" if engine_kind == "pseudo" else "rqmc_full_risk")
    days = [30 * (i + 1) for i in range(12)]
    for g, o in (G.asian(days, [1.0 / 12] * 12, 100.0), G.lookback(days, 95.0, call=False),
                 G.barrier_out(days[-1], days, 100.0, 85.0, down=True)):
        _check_risk(G.graph_request(doc, g, o, days[-1], risk=True))


def test_smooth_graph_greeks():
    """divide, log, exp, negate, the smooth operations and a select."""
    from test_oracle_graph import smooth_graph

    _, doc = G.replay_request("pseudo_mc_full_risk")
    g, out, pay = smooth_graph()
    _check_risk(G.graph_request(doc, g, out, pay, risk=True))


def test_pre_dividend_spot_greeks():
    """Tangents of a pre-dividend spot through its own affine coordinate."""
    _, doc = G.replay_request("pseudo_mc_full_risk")
    doc = copy.deepcopy(doc)
    doc["market"]["discrete_dividends"] = [{"ex_time": 120 / 365, "quote": {"type": "fixed_cash", "amount": 2.0}}]
    g = G.G()
    pre = g.pre_spot(120)
    post = g.spot(120)
    term = g.spot(240)
    ratio = g.push("divide", (post, pre))
    call = g.push("maximum", (g.push("subtract", (term, g.lit(100.0))), g.lit(0.0)))
    out = g.push("add", (g.push("multiply", (call, ratio)), pre))
    _check_risk(G.graph_request(doc, g, out, 240, risk=True))


def test_graph_greeks_in_a_batch_equal_single_plans():
    """A batch with graph products priced and with Greeks beside a fixed
    product: each result bit-identical to its plan evaluated alone."""
    from test_oracle_graph import smooth_graph

    _, doc = G.replay_request("pseudo_mc_full_risk")
    days = [30 * (i + 1) for i in range(6)]
    reqs = [G.graph_request(doc, *G.asian(days, [1.0 / 6] * 6, 100.0), days[-1], risk=True),
            G.graph_request(doc, *G.asian(days, [1.0 / 6] * 6, 100.0), days[-1]),
            G.graph_request(doc, *smooth_graph(), risk=True),
            request.Request.from_v3(copy.deepcopy(doc))]
    plans = [engine.PricingPlan.compile(r) for r in reqs]
    batch = engine.evaluate_batch(plans, "fp64")
    for p, b in zip(plans, batch):
        solo = p.evaluate("fp64")
        for c in range(10):
            assert b.estimates[c].value == solo.estimates[c].value
            assert b.estimates[c].standard_error == solo.estimates[c].standard_error
