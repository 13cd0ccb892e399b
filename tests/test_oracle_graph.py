"""Oracle graph products against the reference's own pins:

* the shared-DAG semantics test of the lazy program (program.rs:259-361):
  100 generated DAGs of Add / Subtract / Maximum / Select over three
  observations, each output equal to plain eager evaluation;
* lazy error semantics (tape.rs:352-470, 856-865): a non-finite value on an
  untaken branch is no error, on an evaluated node it is; a non-finite
  constant fold fails compilation unless the graph branches;
* graph forms of the fixed products reproduce the replay fixtures' prices and
  the oracle's fixed-product prices bit for bit (the reference prices those
  products through the same tape, compile.rs:35-474);
* the product_graph fixture documents against closed forms."""
import ctypes as C
import math

import pytest
from scipy import integrate, stats

import _graphs as G
from _support import oracle, oracle_evaluate
from pricing_amd import abi, request


def _eval(g, output, observed, pre=None, has_pre=None, payment=1000, days=(10, 20, 30)):
    lib = oracle()
    req = request.Request()
    req.set_graph_product(g.nodes, [(output, payment)], [G.U])
    n = len(observed)
    post = (C.c_double * n)(*observed)
    pre_a = (C.c_double * n)(*(pre or [0.0] * n))
    has = (C.c_char * n)(*[bytes([1 if h else 0]) for h in (has_pre or [0] * n)])
    out = C.c_double()
    lib.oracle_graph_evaluate.argtypes = [C.POINTER(abi.pe_product), C.c_uint32, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_char_p,
                                          C.POINTER(C.c_double)]
    day_a = (C.c_int32 * n)(*days[:n])
    rc = lib.oracle_graph_evaluate(C.byref(req.c.product), n, day_a, post, pre_a, C.cast(has, C.c_char_p),
                                   C.byref(out))
    return rc, out.value


def test_generated_shared_dags_match_eager_values():
    for ref, outputs in G.lcg_dags():
        for scenario in range(20):
            observed = [0.0 if scenario % 2 == 0 else 1.0, (scenario % 3) - 1.0, scenario - 10.0]
            values = G.dag_eager(ref, observed)
            for out in sorted(set(outputs)):
                g, o = G.dag_nodes(ref, out)
                rc, v = _eval(g, o, observed)
                assert rc == 0, oracle().oracle_last_error()
                assert v == values[out]


def _log_under_select(threshold):
    g = G.G()
    s = g.spot(10)
    c = g.push("compare", (s, g.lit(threshold)), operation="greater")
    lg = g.push("log", (s,))
    return g, g.push("select", (c, lg, g.lit(0.0)))


def test_untaken_non_finite_branch_is_no_error():
    g, out = _log_under_select(0.0)
    assert _eval(g, out, [-1.0]) == (0, 0.0)          # log(-1) never evaluated
    rc, v = _eval(g, out, [2.0])
    assert rc == 0 and v == math.log(2.0)


def test_evaluated_non_finite_value_is_an_error():
    g = G.G()
    s = g.spot(10)
    out = g.push("log", (s,))
    assert _eval(g, out, [-1.0])[0] == 3
    # masked by a maximum, still evaluated: the reference fails on the log itself
    g2 = G.G()
    s2 = g2.spot(10)
    m = g2.push("maximum", (g2.push("log", (s2,)), g2.lit(0.0)))
    assert _eval(g2, m, [-1.0])[0] == 3


def test_non_finite_constant_fold():
    g = G.G()
    s = g.spot(10)
    bad = g.push("log", (g.lit(-1.0),))
    out = g.push("add", (s, bad))
    assert _eval(g, out, [1.0])[0] == 2                # no Select: compile error
    g2 = G.G()
    s2 = g2.spot(10)
    bad2 = g2.push("log", (g2.lit(-1.0),))
    c = g2.push("compare", (s2, g2.lit(0.0)), operation="less")
    out2 = g2.push("select", (c, bad2, s2))
    assert _eval(g2, out2, [1.0]) == (0, 1.0)          # branching: folded lazily
    assert _eval(g2, out2, [-1.0])[0] == 3


def test_missing_pre_dividend_spot_only_when_evaluated():
    g = G.G()
    s = g.spot(10)
    p = g.pre_spot(10)
    c = g.push("compare", (s, g.lit(0.0)), operation="less")
    out = g.push("select", (c, p, s))
    assert _eval(g, out, [5.0]) == (0, 5.0)
    assert _eval(g, out, [-5.0])[0] == 3
    assert _eval(g, out, [-5.0], pre=[-4.0], has_pre=[1]) == (0, -4.0)


def test_cycles_duplicates_and_late_payments_are_rejected():
    g = G.G()
    g.nodes.append((0, "add", [1, 1], {}))
    g.nodes.append((1, "add", [0, 0], {}))
    assert _eval(g, 0, [])[0] == 2
    g2 = G.G()
    s = g2.spot(30)
    assert _eval(g2, s, [1.0], payment=20)[0] == 2   # depends on an observation after its payment


def fixture_graph(doc):
    """The graph the reference compiles for a European or digital fixture product."""
    pr = doc["product"]
    val = G.dt.date.fromisoformat(doc["valuation_date"])
    expiry = (G.dt.date.fromisoformat(pr["expiry"]) - val).days
    pay = (G.dt.date.fromisoformat(pr.get("payment_date", pr["expiry"])) - val).days
    call = pr["side"]["type"] == "call"
    if pr["type"] == "european_vanilla":
        return G.european(expiry, pr["strike"], pr["notional"], call) + (pay,)
    return G.digital(expiry, pr["strike"], pr["payout"], call, pr["payout_kind"]["type"] == "asset") + (pay,)


BS_FIXTURES = ["pseudo_mc_full_risk", "rqmc_full_risk", "digital_exact_price_only"]


@pytest.mark.parametrize("name", BS_FIXTURES)
def test_graph_product_reproduces_the_replay_fixture(name):
    case, doc = G.replay_request(name)
    g, out, pay = fixture_graph(doc)
    res = oracle_evaluate(G.graph_request(doc, g, out, pay), workers=4, block=case["reduction_block_size"])
    exp = case["value"]
    assert abs(res.estimates[0].value - exp["value"]) <= 1e-12 * abs(exp["value"])
    assert abs(res.estimates[0].standard_error - exp["standard_error"]) <= 1e-10 * exp["standard_error"]


def _fixed_vs_graph(doc, fixed, g, out, pay):
    plain = request.Request.from_v3(G.copy.deepcopy(doc))
    plain.c.risk.delta = plain.c.risk.vega = plain.c.risk.gamma = 0
    fixed(plain)
    a = oracle_evaluate(plain, workers=4)
    b = oracle_evaluate(G.graph_request(doc, g, out, pay), workers=4)
    assert a.estimates[0].value == b.estimates[0].value
    assert a.estimates[0].standard_error == b.estimates[0].standard_error


def test_graph_asian_lookback_and_barrier_equal_the_fixed_products():
    _, doc = G.replay_request("pseudo_mc_full_risk")
    days = [30 * (i + 1) for i in range(6)]
    w = [1.0 / 6] * 6
    g, o = G.asian(days, w, 100.0)
    _fixed_vs_graph(doc, lambda r: r.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, days[-1], 100.0, dates=days,
                                                 weights=w), g, o, days[-1])
    g, o = G.lookback(days, 95.0, call=False)
    _fixed_vs_graph(doc, lambda r: r.set_product(abi.PE_PRODUCT_LOOKBACK, abi.PE_PUT, days[-1], 95.0, dates=days),
                    g, o, days[-1])
    g, o = G.barrier_out(days[-1], days, 100.0, 85.0, down=True)
    _fixed_vs_graph(doc, lambda r: r.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, days[-1], 100.0, barrier=85.0,
                                                 direction_down=1, dates=days), g, o, days[-1])


def _bs_inputs(doc):
    m = doc["market"]
    spot = m["spot"]
    vol = doc["model"]["volatility"]
    return spot, vol


def test_digital_document_matches_its_closed_form():
    case, doc = G.replay_request("pseudo_mc_full_risk")
    pg = G.golden_graph("digital.json")
    val = G.dt.date.fromisoformat(doc["valuation_date"])
    req = G.graph_request(doc, G.G(), 0, 0)
    req.set_product_graph_document(pg, val)
    res = oracle_evaluate(req, workers=4)
    spot, vol = _bs_inputs(doc)
    plain = request.Request.from_v3(G.copy.deepcopy(doc))
    t = (G.dt.date.fromisoformat("2027-09-04") - val).days / 365.0
    pay = (G.dt.date.fromisoformat("2027-09-06") - val).days / 365.0
    fwd, df = _forward_and_discount(doc, t), _discount(doc, pay)
    d2 = (math.log(fwd / 100.0) - 0.5 * vol * vol * t) / (vol * math.sqrt(t))
    exact = df * stats.norm.cdf(d2)
    est = res.estimates[0]
    assert abs(est.value - exact) < 4.0 * est.standard_error, (est.value, exact, est.standard_error)


def _log_df(curve, t):
    # log-linear discount factors between pillars (curve.rs), flat beyond the last pillar's rate
    times, dfs = curve["times"], curve["discount_factors"]
    logs = [math.log(d) for d in dfs]
    if t <= times[0]:
        return logs[0] * t / times[0] if times[0] > 0 else logs[0]
    for i in range(len(times) - 1):
        if t <= times[i + 1]:
            w = (t - times[i]) / (times[i + 1] - times[i])
            return logs[i] + w * (logs[i + 1] - logs[i])
    return logs[-1] * t / times[-1]


def _discount(doc, t):
    return math.exp(_log_df(doc["market"]["discount_curve"], t))


def _forward_and_discount(doc, t):
    m = doc["market"]
    return m["spot"] * math.exp(_log_df(m["dividend_curve"], t) - _log_df(m["discount_curve"], t))


def test_coupon_and_capped_call_document_matches_quadrature():
    _, doc = G.replay_request("pseudo_mc_full_risk")
    pg = G.golden_graph("coupon_and_capped_call.golden.json")
    val = G.dt.date.fromisoformat(doc["valuation_date"])
    req = G.graph_request(doc, G.G(), 0, 0)
    req.set_product_graph_document(pg, val)
    res = oracle_evaluate(req, workers=4)
    spot, vol = _bs_inputs(doc)
    t1 = (G.dt.date.fromisoformat("2027-03-04") - val).days / 365.0
    t2 = (G.dt.date.fromisoformat("2027-09-04") - val).days / 365.0
    p1 = (G.dt.date.fromisoformat("2027-03-06") - val).days / 365.0
    p2 = (G.dt.date.fromisoformat("2027-09-06") - val).days / 365.0
    f1, f2 = _forward_and_discount(doc, t1), _forward_and_discount(doc, t2)

    def smooth_indicator(x, h=2.0):
        if x <= -h:
            return 0.0
        if x >= h:
            return 1.0
        u = (x + h) / (2 * h)
        return u ** 3 * (10 - 15 * u + 6 * u * u)

    sd1 = vol * math.sqrt(t1)
    coupon = integrate.quad(lambda z: smooth_indicator(f1 * math.exp(-0.5 * sd1 * sd1 + sd1 * z) - 95.0)
                            * stats.norm.pdf(z), -12, 12, points=[(math.log(95.0 / f1) + 0.5 * sd1 * sd1) / sd1],
                            limit=200)[0]

    def call(f, k, t):
        s = vol * math.sqrt(t)
        d1 = (math.log(f / k) + 0.5 * s * s) / s
        return f * stats.norm.cdf(d1) - k * stats.norm.cdf(d1 - s)

    spread = call(f2, 100.0, t2) - call(f2, 120.0, t2)
    exact = 5.0 * coupon * _discount(doc, p1) + spread * _discount(doc, p2)
    est = res.estimates[0]
    assert abs(est.value - exact) < 4.0 * est.standard_error, (est.value, exact, est.standard_error)


# ---- graph products on local-vol paths (local_vol_valuation.rs:221-247) ----

LV_FIXTURES = ["pseudo_mc_price_only", "rqmc_price_only"]


@pytest.mark.parametrize("name", LV_FIXTURES)
def test_graph_european_under_local_vol_reproduces_the_replay_fixture(name):
    case, doc = G.replay_request(name, "local_volatility")
    g, out, pay = fixture_graph(doc)
    res = oracle_evaluate(G.graph_request(doc, g, out, pay), workers=4, block=case["reduction_block_size"])
    exp = case["value"]
    assert abs(res.estimates[0].value - exp["value"]) <= 1e-12 * abs(exp["value"])
    assert abs(res.estimates[0].standard_error - exp["standard_error"]) <= 1e-10 * exp["standard_error"]


def test_graph_path_dependence_under_local_vol_equals_the_fixed_products():
    """Monthly Asian, lookback and barrier graphs ending at the local-vol
    grid's expiry (day 365) price exactly as the fixed products on the same
    local-vol paths."""
    _, doc = G.replay_request("pseudo_mc_price_only", "local_volatility")
    days = [30 * (i + 1) for i in range(12)] + [365]
    w = [1.0 / 13] * 13
    g, o = G.asian(days, w, 100.0)
    _fixed_vs_graph(doc, lambda r: r.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, days[-1], 100.0, dates=days,
                                                 weights=w), g, o, days[-1])
    g, o = G.lookback(days, 95.0, call=False)
    _fixed_vs_graph(doc, lambda r: r.set_product(abi.PE_PRODUCT_LOOKBACK, abi.PE_PUT, days[-1], 95.0, dates=days),
                    g, o, days[-1])
    g, o = G.barrier_out(days[-1], days, 100.0, 85.0, down=True)
    _fixed_vs_graph(doc, lambda r: r.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, days[-1], 100.0, barrier=85.0,
                                                 direction_down=1, dates=days), g, o, days[-1])


def test_pre_dividend_spot_under_local_vol_is_refused():
    """Local-vol paths carry no discrete dividends here: a graph reading a
    pre-dividend spot is refused at compile time (oracle and engine)."""
    from pricing_amd import engine
    _, doc = G.replay_request("pseudo_mc_price_only", "local_volatility")
    g = G.G()
    s = g.pre_spot(365)
    k = g.lit(100.0)
    out = g.push("maximum", (g.push("subtract", (s, k)), g.lit(0.0)))
    req = G.graph_request(doc, g, out, 365)
    with pytest.raises(RuntimeError):
        oracle_evaluate(req, workers=1)
    with pytest.raises(engine.EngineError) as err:
        engine.PricingPlan.compile(req)
    assert err.value.status == abi.PE_ERR_UNSUPPORTED


# ---- graph product Greeks: the reverse sweep (tape.rs:484-805) ----

RISK_FIXTURES = ["pseudo_mc_full_risk", "rqmc_full_risk"]


@pytest.mark.parametrize("name", RISK_FIXTURES)
def test_graph_greeks_reproduce_the_replay_fixture(name):
    """The European graph's pathwise Delta and Vega, and its Gamma from
    Deltas at the bumped spots, reproduce the fixture's Greeks."""
    case, doc = G.replay_request(name)
    g, out, pay = fixture_graph(doc)
    res = oracle_evaluate(G.graph_request(doc, g, out, pay, risk=True), workers=4, block=case["reduction_block_size"])
    for key, index in (("delta", 1), ("vega", 2), ("gamma", 3)):
        e = case["risks"][key]
        got = res.estimates[index]
        assert abs(got.value - e["value"]) <= 1e-12 * abs(e["value"]), key
        assert abs(got.standard_error - e["standard_error"]) <= 1e-10 * e["standard_error"], key


def _fixed_vs_graph_risk(doc, fixed, g, out, pay):
    plain = request.Request.from_v3(G.copy.deepcopy(doc))
    fixed(plain)
    a = oracle_evaluate(plain, workers=4)
    b = oracle_evaluate(G.graph_request(doc, g, out, pay, risk=True), workers=4)
    for c in range(10):
        x, y = a.estimates[c], b.estimates[c]
        scale = max(abs(x.value), x.standard_error, 1e-300)
        assert abs(x.value - y.value) <= 1e-12 * scale, (c, x.value, y.value)
        assert abs(x.standard_error - y.standard_error) <= 1e-10 * max(x.standard_error, 1e-300), c


def test_graph_greeks_of_asian_lookback_and_barrier_equal_the_fixed_products():
    """Every pullback the Asian, lookback and barrier graphs use (multiply,
    add, subtract, maximum, minimum, indicator, barrier-hit state) against the
    fixed products' own tapes, on the same paths."""
    _, doc = G.replay_request("pseudo_mc_full_risk")
    days = [30 * (i + 1) for i in range(6)]
    w = [1.0 / 6] * 6
    g, o = G.asian(days, w, 100.0)
    _fixed_vs_graph_risk(doc, lambda r: r.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, days[-1], 100.0, dates=days,
                                                      weights=w), g, o, days[-1])
    g, o = G.lookback(days, 95.0, call=False)
    _fixed_vs_graph_risk(doc, lambda r: r.set_product(abi.PE_PRODUCT_LOOKBACK, abi.PE_PUT, days[-1], 95.0,
                                                      dates=days), g, o, days[-1])
    g, o = G.barrier_out(days[-1], days, 100.0, 85.0, down=True)
    _fixed_vs_graph_risk(doc, lambda r: r.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, days[-1], 100.0,
                                                      barrier=85.0, direction_down=1, dates=days), g, o, days[-1])


def smooth_graph(d1=90, d2=180, k=100.0):
    """A payoff smooth in both spots that uses the pullbacks no fixed product
    has: divide, log, exp, negate, the three smooth operations and a select on
    a comparison of two smooth quantities (always true here)."""
    g = G.G()
    a, b = g.spot(d1), g.spot(d2)
    ratio = g.push("divide", (b, a))
    inv = g.push("exp", (g.push("negate", (g.push("log", (ratio,)),)),))          # a / b
    call = g.push("smooth_maximum", (g.push("subtract", (b, g.lit(k))), g.lit(0.0)), half_width=2.0)
    cap = g.push("smooth_minimum", (b, g.push("multiply", (a, g.lit(1.1)))), half_width=3.0)
    dig = g.push("smooth_indicator", (g.push("subtract", (b, g.lit(k))),), half_width=5.0)
    body = g.push("add", (g.push("multiply", (call, inv)), g.push("add", (cap, g.push("multiply", (dig, g.lit(10.0)))))))
    cond = g.push("compare", (g.push("add", (cap, g.lit(1e6))), b), operation="greater")
    return g, g.push("select", (cond, body, g.lit(-1.0))), d2


def test_smooth_graph_greeks_agree_with_bumped_prices():
    """For a payoff smooth in the spots the pathwise Greeks equal the
    central differences of the price on the same paths up to the bump's
    O(h^2) term: a wrong pullback (sign, dropped factor, swapped operands)
    of any operation in the graph moves Delta, Gamma or Vega by far more."""
    _, doc = G.replay_request("pseudo_mc_full_risk")
    g, out, pay = smooth_graph()
    res = oracle_evaluate(G.graph_request(doc, g, out, pay, risk=True), workers=4)
    e = res.estimates
    # relative gaps measured: Delta 1e-5, Vega 1e-7, Gamma 2.5e-3 (Gamma's
    # two differences carry different O(h^2) terms)
    for greek, bumped, tol in ((1, 4, 1e-4), (2, 6, 1e-5), (3, 8, 1e-2)):
        assert abs(e[greek].value - e[bumped].value) <= tol * abs(e[greek].value), (greek, e[greek].value,
                                                                                  e[bumped].value)
