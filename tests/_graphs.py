"""Graph-product builders for the tests: the reference's product graphs
(crates/pricing/src/engine/payoff/compile.rs:35-474) written as
``Request.set_graph_product`` node lists, and the generator of the
reference's own shared-DAG semantics test (program.rs:259-361)."""
import copy
import datetime as dt
import json
import os

from _support import GOLDEN, golden
from pricing_amd import request

U = 1  # the fixtures' underlying id


class G:
    """Node-list builder; ids increase from 0."""

    def __init__(self):
        self.nodes = []

    def push(self, op, operands=(), **extra):
        nid = len(self.nodes)
        self.nodes.append((nid, op, list(operands), extra))
        return nid

    def lit(self, v):
        return self.push("literal", value=float(v))

    def spot(self, day):
        return self.push("terminal_spot", underlying_id=U, observation_day=day)

    def pre_spot(self, day):
        return self.push("pre_dividend_spot", underlying_id=U, observation_day=day)


def european(expiry, strike, notional=1.0, call=True):
    g = G()
    s = g.spot(expiry)
    k = g.lit(strike)
    d = g.push("subtract", (s, k) if call else (k, s))
    z = g.lit(0.0)
    pp = g.push("maximum", (d, z))
    n = g.lit(notional)
    return g, g.push("multiply", (pp, n))


def digital(expiry, strike, payout, call=True, asset=False):
    g = G()
    s = g.spot(expiry)
    k = g.lit(strike)
    d = g.push("subtract", (s, k) if call else (k, s))
    ind = g.push("indicator", (d,))
    pay = g.lit(payout)
    base = g.push("multiply", (s, pay)) if asset else pay
    return g, g.push("multiply", (ind, base))


def asian(days, weights, strike, notional=1.0, call=True):
    g = G()
    ws = g.lit(0.0)
    for d, w in zip(days, weights):
        o = g.spot(d)
        wl = g.lit(w)
        ws = g.push("add", (ws, g.push("multiply", (o, wl))))
    k = g.lit(strike)
    d = g.push("subtract", (ws, k) if call else (k, ws))
    z = g.lit(0.0)
    pp = g.push("maximum", (d, z))
    return g, g.push("multiply", (pp, g.lit(notional)))


def lookback(days, strike, notional=1.0, call=True):
    g = G()
    ext = g.spot(days[0])
    for d in days[1:]:
        ext = g.push("maximum" if call else "minimum", (ext, g.spot(d)))
    k = g.lit(strike)
    d = g.push("subtract", (ext, k) if call else (k, ext))
    pp = g.push("maximum", (d, g.lit(0.0)))
    return g, g.push("multiply", (pp, g.lit(notional)))


def barrier_out(expiry, days, strike, barrier, down, notional=1.0, call=True):
    """Discrete knock-out (compile.rs:280-402 without rebate or smoothing)."""
    g = G()
    k = g.lit(strike)
    term = g.spot(expiry)
    sd = g.push("subtract", (term, k) if call else (k, term))
    z = g.lit(0.0)
    pp = g.push("maximum", (sd, z))
    vanilla = g.push("multiply", (pp, g.lit(notional)))
    b = g.lit(barrier)
    hit = z
    for d in days:
        s = g.spot(d)
        dist = g.push("subtract", (b, s) if down else (s, b))
        hit = g.push("barrier_hit_state", (hit, g.push("indicator", (dist,))))
    active = g.push("subtract", (g.lit(1.0), hit))
    return g, g.push("multiply", (vanilla, active))


def replay_request(name, source=""):
    """A Black-Scholes replay fixture's request document (market, model, engine)."""
    case = next(c for c in golden()["replay"] if c["name"] == name and source in c["source"])
    return case, copy.deepcopy(case["request"])


def graph_request(doc, g, output, payment_day, risk=False):
    """`doc`'s market, model and engine with the graph as the product; with
    `risk`, also `doc`'s Greek requests."""
    plain = copy.deepcopy(doc)
    if not risk:
        plain["risk"] = {"delta": False, "vega": False}
    req = request.Request.from_v3(plain)
    req.set_graph_product(g.nodes, [(output, payment_day)], [U])
    return req


def golden_graph(name):
    with open(os.path.join(GOLDEN, "product_graphs", name)) as f:
        return json.load(f)


def lcg_dags(count=100, ops=50, seed=1847):
    """The reference's generated shared DAGs (program.rs:261-325): three
    observations, then `ops` random Add / Subtract / Maximum / Select nodes
    over earlier nodes; returns (nodes as (kind, a, b, c), outputs)."""
    state = [seed]

    def nxt():
        state[0] = (state[0] * 6364136223846793005 + 1) % (1 << 64)
        return state[0] >> 32

    graphs = []
    for _ in range(count):
        ref = [("obs", i, 0, 0) for i in range(3)]
        for _ in range(ops):
            n = len(ref)
            left, right, cond = nxt() % n, nxt() % n, nxt() % 3
            kind = ("add", "subtract", "maximum", "select")[nxt() % 4]
            ref.append((kind, left, right, cond))
        graphs.append((ref, [len(ref) - 1, len(ref) - 3, len(ref) - 1]))
    return graphs


def dag_nodes(ref, output, days=(10, 20, 30)):
    g = G()
    for kind, a, b, c in ref:
        if kind == "obs":
            g.spot(days[a])
        elif kind == "select":
            g.push("select", (c, a, b))
        else:
            g.push(kind, (a, b))
    return g, output


def dag_eager(ref, observed):
    v = []
    for kind, a, b, c in ref:
        if kind == "obs":
            v.append(observed[a])
        elif kind == "add":
            v.append(v[a] + v[b])
        elif kind == "subtract":
            v.append(v[a] - v[b])
        elif kind == "maximum":
            v.append(v[a] if v[a] >= v[b] else v[b])
        else:
            v.append(v[a] if v[c] != 0.0 else v[b])
    return v


VALUATION = dt.date(2026, 9, 4)
