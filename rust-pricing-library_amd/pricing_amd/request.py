"""PricingRequest mirror: builds the C-ABI request from the reference's
schema-v3 request documents.

Field names and meaning follow ``schemas/v3/pricing_request.schema.json`` and
``PricingRequest`` (crates/pricing/src/api/request.rs:11-300).  Dates become
whole days from the valuation date (ACT/365F, crates/pricing/src/core/date.rs:223-231).
Only the single-asset products and models of the simulation hot path are
accepted; anything else raises ``UnsupportedRequest``.
"""
import ctypes as C
import datetime as _dt
import math

from . import abi


class UnsupportedRequest(ValueError):
    pass


def _days(valuation, date):
    return (_dt.date.fromisoformat(date) - valuation).days


def _darr(values):
    arr = (C.c_double * len(values))(*values)
    return arr


def _iarr(values):
    return (C.c_int32 * len(values))(*values)


class Request:
    """A C-ABI request plus the buffers it points into (kept alive here)."""

    def __init__(self):
        self.c = abi.pe_request()
        self._keep = []
        self.doc = None

    def _ptr_d(self, values):
        arr = _darr(values)
        self._keep.append(arr)
        return C.cast(arr, C.POINTER(C.c_double))

    def _ptr_i(self, values):
        arr = _iarr(values)
        self._keep.append(arr)
        return C.cast(arr, C.POINTER(C.c_int32))

    def pointer(self):
        return C.pointer(self.c)

    # -------------------------------------------------------------- builders
    def set_market(self, spot, discount_times, discount_factors, dividend_times, dividend_factors,
                   dividends=()):
        m = self.c.market
        m.spot = float(spot)
        m.discount_curve.count = len(discount_times)
        m.discount_curve.times = self._ptr_d(discount_times)
        m.discount_curve.discount_factors = self._ptr_d(discount_factors)
        m.dividend_curve.count = len(dividend_times)
        m.dividend_curve.times = self._ptr_d(dividend_times)
        m.dividend_curve.discount_factors = self._ptr_d(dividend_factors)
        if dividends:
            arr = (abi.pe_dividend_quote * len(dividends))()
            for i, (ex_time, fixed_cash, beta) in enumerate(dividends):
                arr[i].ex_time = float(ex_time)
                arr[i].fixed_cash = float(fixed_cash)
                arr[i].beta = float(beta)
            self._keep.append(arr)
            m.dividend_count = len(dividends)
            m.dividends = C.cast(arr, C.POINTER(abi.pe_dividend_quote))
        return self

    def set_black_scholes(self, volatility, black76=False):
        self.c.model.kind = abi.PE_MODEL_BLACK76 if black76 else abi.PE_MODEL_BLACK_SCHOLES
        self.c.model.volatility = float(volatility)
        return self

    def set_local_vol(self, time_nodes, x_nodes, values, floor, cap):
        self.c.model.kind = abi.PE_MODEL_LOCAL_VOL
        g = self.c.model.lv
        g.time_count = len(time_nodes)
        g.x_count = len(x_nodes)
        g.time_nodes = self._ptr_d(time_nodes)
        g.x_nodes = self._ptr_d(x_nodes)
        g.values = self._ptr_d(values)
        g.floor = float(floor)
        g.cap = float(cap)
        return self

    def set_product(self, kind, side, expiry_day, strike, notional=1.0, payment_day=None, **kw):
        p = self.c.product
        p.kind = kind
        p.side = side
        p.expiry_day = int(expiry_day)
        p.payment_day = int(expiry_day if payment_day is None else payment_day)
        p.strike = float(strike)
        p.notional = float(notional)
        p.payout = float(kw.get("payout", 0.0))
        p.payout_is_asset = int(kw.get("payout_is_asset", 0))
        p.barrier = float(kw.get("barrier", 0.0))
        p.direction_down = int(kw.get("direction_down", 0))
        p.knock_in = int(kw.get("knock_in", 0))
        p.continuous = int(kw.get("continuous", 0))
        rebate = kw.get("rebate")
        p.has_rebate = 0 if rebate is None else 1
        p.rebate = 0.0 if rebate is None else float(rebate)
        hh = kw.get("historical_hit")
        p.historical_hit = -1 if hh is None else int(bool(hh))
        he = kw.get("historical_extremum")
        p.has_historical_extremum = 0 if he is None else 1
        p.historical_extremum = 0.0 if he is None else float(he)
        dates = kw.get("dates", ())
        p.date_count = len(dates)
        p.dates = self._ptr_i(dates) if dates else None
        weights = kw.get("weights")
        p.weights = self._ptr_d(weights) if weights else None
        fixings = kw.get("fixings")
        p.fixings = self._ptr_d([math.nan if f is None else f for f in fixings]) if fixings else None
        return self

    def set_graph_product(self, nodes, cashflows, underlyings):
        """GraphProduct (crates/pricing/src/product/graph_product.rs:40-137).

        nodes: (id, operation name, operands, extra) with operands the node ids
        in the order of ``pe_graph_opcode`` and extra a dict holding
        ``value`` (literal), ``half_width`` (smooth ops), ``operation``
        (compare), ``underlying_id`` / ``observation_day`` (spots).
        cashflows: (output node id, payment day); underlyings: declared ids."""
        p = self.c.product
        p.kind = abi.PE_PRODUCT_GRAPH
        arr = (abi.pe_graph_node * len(nodes))()
        for i, (nid, op, operands, extra) in enumerate(nodes):
            n = arr[i]
            n.id = int(nid)
            if op not in abi.GRAPH_OPCODES:
                raise UnsupportedRequest(f"graph operation {op}")
            n.opcode = abi.GRAPH_OPCODES[op]
            for j, o in enumerate(operands):
                n.operands[j] = int(o)
            n.comparison = abi.GRAPH_COMPARISONS.get(extra.get("operation", "less"), 0)
            n.underlying_id = int(extra.get("underlying_id", 0))
            n.observation_day = int(extra.get("observation_day", 0))
            n.value = float(extra.get("value", extra.get("half_width", 0.0)))
        flows = (abi.pe_graph_cashflow * len(cashflows))()
        for i, (out, day) in enumerate(cashflows):
            flows[i].output_node = int(out)
            flows[i].payment_day = int(day)
        unders = (C.c_uint32 * len(underlyings))(*[int(u) for u in underlyings])
        self._keep.extend([arr, flows, unders])
        p.graph_node_count = len(nodes)
        p.graph_nodes = C.cast(arr, C.POINTER(abi.pe_graph_node))
        p.graph_cashflow_count = len(cashflows)
        p.graph_cashflows = C.cast(flows, C.POINTER(abi.pe_graph_cashflow))
        p.graph_underlying_count = len(underlyings)
        p.graph_underlyings = C.cast(unders, C.POINTER(C.c_uint32))
        return self

    def set_product_graph_document(self, doc, valuation):
        """A ``product_graph`` document (schemas/product-graph-v1|v2/
        product_graph.schema.json) as the product; dates become days after
        `valuation` (a datetime.date)."""
        if doc.get("document_kind") != "product_graph":
            raise UnsupportedRequest("not a product_graph document")
        prod = doc["product"]
        operand_keys = ("left", "right", "input", "numerator", "denominator", "previous", "hit_weight",
                        "condition", "then_value", "else_value")
        nodes = []
        for n in prod["nodes"]:
            op = dict(n["operation"])
            kind = op.pop("type")
            operands = [op.pop(k) for k in operand_keys if k in op]
            if kind == "divide":
                operands = [n["operation"]["numerator"], n["operation"]["denominator"]]
            elif kind == "barrier_hit_state":
                operands = [n["operation"]["previous"], n["operation"]["hit_weight"]]
            elif kind == "select":
                operands = [n["operation"][k] for k in ("condition", "then_value", "else_value")]
            if "observation_date" in op:
                op["observation_day"] = _days(valuation, op.pop("observation_date"))
            nodes.append((n["id"], kind, operands, op))
        flows = [(c["output_node"], _days(valuation, c["payment_date"])) for c in prod["cashflows"]]
        return self.set_graph_product(nodes, flows, prod["underlying_ids"])

    def set_pseudo(self, seed, units, antithetic=False, brownian_bridge=False):
        e = self.c.engine
        e.kind = abi.PE_ENGINE_PSEUDO
        e.seed = int(seed)
        e.units = int(units)
        e.scrambles = 0
        e.antithetic = int(bool(antithetic))
        e.brownian_bridge = int(bool(brownian_bridge))
        return self

    def set_rqmc(self, seed, points, scrambles, antithetic=False, brownian_bridge=False):
        e = self.c.engine
        e.kind = abi.PE_ENGINE_RQMC
        e.seed = int(seed)
        e.units = int(points)
        e.scrambles = int(scrambles)
        e.antithetic = int(bool(antithetic))
        e.brownian_bridge = int(bool(brownian_bridge))
        return self

    def set_lsm(self, training, antithetic=False, brownian_bridge=False, max_degree=3, abs_rank_tolerance=1e-14,
                rel_rank_tolerance=1e-12, itm_abs_tolerance=0.0, max_matrix_elements=1 << 40,
                method=abi.PE_LSM_POLY_CPQR, mlp_hidden=64, mlp_epochs=4, mlp_batch=8192, mlp_learning_rate=1e-3,
                mlp_seed=0x4D4C50):
        """LsmConfig: training = (engine kind, seed, units or points, scrambles)."""
        l = self.c.lsm
        kind, seed, units, scrambles = training
        l.training_engine.kind = kind
        l.training_engine.seed = int(seed)
        l.training_engine.units = int(units)
        l.training_engine.scrambles = int(scrambles)
        l.training_engine.antithetic = int(bool(antithetic))
        l.training_engine.brownian_bridge = int(bool(brownian_bridge))
        l.max_degree = int(max_degree)
        l.abs_rank_tolerance = float(abs_rank_tolerance)
        l.rel_rank_tolerance = float(rel_rank_tolerance)
        l.itm_abs_tolerance = float(itm_abs_tolerance)
        l.max_matrix_elements = int(max_matrix_elements)
        l.method = int(method)
        l.mlp_hidden = int(mlp_hidden)
        l.mlp_epochs = int(mlp_epochs)
        l.mlp_batch = int(mlp_batch)
        l.mlp_learning_rate = float(mlp_learning_rate)
        l.mlp_seed = int(mlp_seed)
        return self

    def set_risk(self, delta=False, vega=False, gamma_bump=None, gamma_relative=True,
                 smoothing_half_width=None, aad_tile_capacity=None, checkpoint_interval=None):
        r = self.c.risk
        r.aad_tile_capacity = int(aad_tile_capacity or 0)
        r.checkpoint_interval = int(checkpoint_interval or 0)
        r.delta = int(bool(delta))
        r.vega = int(bool(vega))
        r.gamma = 0 if gamma_bump is None else 1
        r.gamma_relative = int(bool(gamma_relative))
        r.gamma_bump = 0.0 if gamma_bump is None else float(gamma_bump)
        r.smoothing = 0 if smoothing_half_width is None else 1
        r.smoothing_half_width = 0.0 if smoothing_half_width is None else float(smoothing_half_width)
        return self

    # ------------------------------------------------------- schema-v3 reader
    @classmethod
    def from_v3(cls, doc):
        """Builds a request from a schema-v3 ``pricing_request`` document."""
        self = cls()
        self.doc = doc
        valuation = _dt.date.fromisoformat(doc["valuation_date"])
        mk = doc["market"]
        if mk.get("type") != "equity":
            raise UnsupportedRequest("only equity markets are supported")
        dividends = []
        for d in mk.get("discrete_dividends", []) or []:
            # DividendEventV1 (crates/pricing/src/wire.rs:480-492, 1617-1633)
            q = d["quote"]
            if q["type"] == "fixed_cash":
                dividends.append((d["ex_time"], q["amount"], 0.0))
            elif q["type"] == "proportional":
                dividends.append((d["ex_time"], 0.0, q["beta"]))
            elif q["type"] == "fixed_cash_and_proportional":
                dividends.append((d["ex_time"], q["fixed_cash"], q["beta"]))
            else:
                raise UnsupportedRequest(f"dividend quote {q['type']}")
        self.set_market(mk["spot"], mk["discount_curve"]["times"], mk["discount_curve"]["discount_factors"],
                        mk["dividend_curve"]["times"], mk["dividend_curve"]["discount_factors"], dividends)
        model = doc["model"]
        if model["type"] == "black_scholes":
            self.set_black_scholes(model["volatility"])
        elif model["type"] == "black_76":
            self.set_black_scholes(model["volatility"], black76=True)
        elif model["type"] == "local_volatility":
            g = model["local_variance_grid"]
            self.set_local_vol(g["time_nodes"], g["log_forward_moneyness_nodes"], g["values"], g["floor"], g["cap"])
        else:
            raise UnsupportedRequest(f"model {model['type']}")
        pr = doc["product"]
        side = abi.PE_CALL if pr.get("side", {}).get("type", "call") == "call" else abi.PE_PUT
        kind = pr["type"]
        notional = pr.get("notional", 1.0)
        if kind == "european_vanilla":
            self.set_product(abi.PE_PRODUCT_EUROPEAN, side, _days(valuation, pr["expiry"]), pr["strike"], notional,
                             _days(valuation, pr.get("payment_date", pr["expiry"])))
        elif kind == "digital":
            self.set_product(abi.PE_PRODUCT_DIGITAL, side, _days(valuation, pr["expiry"]), pr["strike"], 1.0,
                             _days(valuation, pr.get("payment_date", pr["expiry"])), payout=pr["payout"],
                             payout_is_asset=int(pr["payout_kind"]["type"] == "asset"))
        elif kind == "barrier":
            self.set_product(
                abi.PE_PRODUCT_BARRIER, side, _days(valuation, pr["expiry"]), pr["strike"], notional,
                _days(valuation, pr.get("payment_date", pr["expiry"])), barrier=pr["barrier"],
                direction_down=int(pr["direction"]["type"] == "down"),
                knock_in=int(pr["style"]["type"] == "knock_in"),
                continuous=int(pr["monitoring"]["type"] == "continuous"),
                rebate=pr.get("rebate"), historical_hit=pr.get("historical_hit"),
                dates=[_days(valuation, d) for d in pr["monitoring_dates"]])
        elif kind == "arithmetic_asian":
            obs = pr["observations"]
            last = max(_dt.date.fromisoformat(o["date"]) for o in obs)
            dates = [_days(valuation, o["date"]) for o in obs]
            weights = [o["weight"] for o in obs]
            fixings = [o["value"]["fixing"] if o["value"]["type"] == "known" else None for o in obs]
            self.set_product(abi.PE_PRODUCT_ASIAN, side, (last - valuation).days, pr["strike"], notional,
                             _days(valuation, pr.get("payment_date", last.isoformat())), dates=dates,
                             weights=weights, fixings=fixings)
        elif kind == "american_vanilla":
            dates = [_days(valuation, d) for d in pr["exercise_dates"]]
            self.set_product(abi.PE_PRODUCT_AMERICAN, side, _days(valuation, pr["expiry"]), pr["strike"], notional,
                             _days(valuation, pr.get("payment_date", pr["expiry"])), dates=dates)
        elif kind == "fixed_lookback":
            dates = [_days(valuation, d) for d in pr["monitoring_dates"]]
            self.set_product(abi.PE_PRODUCT_LOOKBACK, side, max(dates), pr["strike"], notional,
                             _days(valuation, pr.get("payment_date", pr["monitoring_dates"][-1])), dates=dates,
                             historical_extremum=pr.get("historical_extremum"))
        else:
            raise UnsupportedRequest(f"product {kind}")
        eng = doc["engine"]
        vr = eng.get("variance_reduction", {})
        if eng["type"] == "pseudo_monte_carlo":
            self.set_pseudo(eng["master_seed"], eng["independent_sampling_units"], vr.get("antithetic", False),
                            vr.get("brownian_bridge", False))
        elif eng["type"] == "randomized_quasi_monte_carlo":
            self.set_rqmc(eng["master_scramble_seed"], eng["points_per_scramble"], eng["scramble_count"],
                          vr.get("antithetic", False), vr.get("brownian_bridge", False))
        else:
            raise UnsupportedRequest(f"engine {eng['type']}")
        lsm = doc.get("lsm")
        if lsm:
            t = lsm["training_engine"]
            tvr = t.get("variance_reduction", {})
            if t["type"] == "pseudo_monte_carlo":
                training = (abi.PE_ENGINE_PSEUDO, t["master_seed"], t["independent_sampling_units"], 0)
            else:
                training = (abi.PE_ENGINE_RQMC, t["master_scramble_seed"], t["points_per_scramble"], t["scramble_count"])
            if [v["type"] for v in lsm["state_variables"]] != ["spot"] or lsm["basis"]["feature_count"] != 1:
                raise UnsupportedRequest("LSM state variables other than the spot")
            self.set_lsm(training, tvr.get("antithetic", False), tvr.get("brownian_bridge", False),
                         lsm["basis"]["max_degree"], lsm["cpqr"]["abs_rank_tolerance"],
                         lsm["cpqr"]["rel_rank_tolerance"], lsm["itm_abs_tolerance"], lsm["max_matrix_elements"])
        risk = doc.get("risk", {})
        gamma = risk.get("gamma")
        gamma_bump = None
        gamma_relative = True
        if gamma:
            gamma_bump = gamma["bump"]["value"]
            gamma_relative = gamma["bump"]["type"] == "relative"
        smoothing = risk.get("payoff_smoothing")
        self.set_risk(risk.get("delta", False), risk.get("vega", False), gamma_bump, gamma_relative,
                      smoothing["half_width"] if smoothing else None, risk.get("aad_tile_capacity"),
                      risk.get("checkpoint_interval"))
        return self


def exec_policy(worker_threads=1, reduction_block_size=0):
    p = abi.pe_exec_policy()
    p.worker_threads = int(worker_threads)
    p.reduction_block_size = int(reduction_block_size)
    return p
