"""Pins the oracle's continuous-barrier bridge, compact-C2 smoothing, dividend
jump monitoring and escrowed affine dividends.

Sources of truth:
* the reference's Decimal fixtures (fixtures/path-dependence/reference-cases-v0.1.json,
  fixtures/local-vol/reference-cases-v0.1.json affine_dividend), extracted into
  tests/golden/reference_vectors.json;
* the frozen replay of the smoothed continuous barrier with full risk, its
  bump-and-revalue validation and its bridge diagnostics
  (fixtures/replay/path_dependence-windows-x86_64.json);
* closed forms: Reiner-Rubinstein up-and-out call for the exact bridge
  (the bridge crossing probability is exact for GBM, so the estimator is
  unbiased), the textbook escrowed-dividend forward and Black-76 price,
  the expected average of an Asian, and a zero-volatility barrier;
* central finite differences of the interval log-survival.
"""
import ctypes as C
import math
import struct

import pytest
from scipy.stats import norm

from _support import darr, golden, oracle, oracle_evaluate
from pricing_amd import abi, request

G = golden()
PD = G["path_dependence"]


def _d(x):
    return float(x)


def smoothing(h, x, y=0.0):
    out = (C.c_double * 8)()
    oracle().oracle_smoothing(h, x, y, out)
    return list(out)


@pytest.mark.parametrize("case", PD["smoothing_cases"], ids=lambda c: c["id"])
def test_smoothing_cases(case):
    h, x = _d(case["half_width"]), _d(case["input"])
    got = smoothing(h, x)
    e = case["expected"]
    keys = ["indicator", "indicator_first", "indicator_second", "positive_part", "positive_part_first",
            "positive_part_second"]
    for k, g in zip(keys, got[:6]):
        assert abs(g - _d(e[k])) <= 1e-15 * max(1.0, abs(_d(e[k]))), k


@pytest.mark.parametrize("case", PD["extrema_cases"], ids=lambda c: c["id"])
def test_smooth_extrema_cases(case):
    got = smoothing(_d(case["half_width"]), _d(case["left"]), _d(case["right"]))
    assert abs(got[6] - _d(case["expected"]["maximum"])) <= 1e-13
    assert abs(got[7] - _d(case["expected"]["minimum"])) <= 1e-13


def bridge_interval(up, ls, rs, lb, rb, lv, rv, dt):
    out = (C.c_double * 9)()
    rc = oracle().oracle_bridge_interval(int(up), ls, rs, lb, rb, lv, rv, dt, out)
    assert rc == 0, oracle().oracle_last_error()
    return list(out)


@pytest.mark.parametrize("case", PD["bridge_cases"], ids=lambda c: c["id"])
def test_exact_bridge_cases(case):
    up = case["direction"] == "up"
    b, s0, s1, iv = (_d(case[k]) for k in ("barrier", "start_spot", "end_spot", "integrated_variance"))
    out = bridge_interval(up, s0, s1, b, b, iv, iv, 1.0)
    survival = math.exp(out[1])
    exp = _d(case["expected"]["survival_probability"])
    # log-space survival: ln then exp costs a few ulps
    assert abs(survival - exp) <= 1e-14 * max(1.0, exp)


def test_exact_bridge_derivatives_match_finite_differences():
    base = dict(ls=90.0, rs=92.0, lb=100.0, rb=101.0, lv=0.04, rv=0.06, dt=0.25)
    names = ["ls", "rs", "lb", "rb", "lv", "rv", "dt"]
    out = bridge_interval(True, *base.values())
    assert out[0] == 0.0  # finite correction
    for k, name in enumerate(names):
        h = 1e-6 * abs(base[name])
        up = dict(base); up[name] += h
        dn = dict(base); dn[name] -= h
        fd = (bridge_interval(True, *up.values())[1] - bridge_interval(True, *dn.values())[1]) / (2 * h)
        assert abs(out[2 + k] - fd) <= 1e-6 * max(1.0, abs(fd)), name


def test_exact_bridge_down_mirror():
    # A down barrier on (S, B) is the up barrier on (1/S, 1/B).
    a = bridge_interval(False, 100.0, 90.0, 80.0, 80.0, 0.04, 0.04, 1.0)
    b = bridge_interval(True, 1 / 100.0, 1 / 90.0, 1 / 80.0, 1 / 80.0, 0.04, 0.04, 1.0)
    assert abs(a[1] - b[1]) <= 1e-13


def smoothed_bridge(up, s0, s1, b, h, iv):
    out = (C.c_double * 12)()
    rc = oracle().oracle_smoothed_bridge(int(up), s0, s1, b, h, iv, out)
    assert rc == 0, oracle().oracle_last_error()
    return list(out)


@pytest.mark.parametrize("case", PD["smoothed_bridge_cases"], ids=lambda c: c["id"])
def test_smoothed_bridge_cases(case):
    up = case["direction"] == "up"
    got = smoothed_bridge(up, _d(case["start_spot"]), _d(case["end_spot"]), _d(case["barrier"]),
                          _d(case["half_width"]), _d(case["integrated_variance"]))
    keys = ["start_hit_weight", "end_hit_weight", "start_effective_log_distance", "end_effective_log_distance",
            "bridge_survival", "endpoint_survival", "total_survival", "total_hit_weight"]
    for k, g in zip(keys, got[:8]):
        e = _d(case["expected"][k])
        assert abs(g - e) <= 2e-15 * max(1.0, abs(e)), k


def test_smoothed_bridge_variance_derivative_matches_finite_difference():
    h = 1e-7
    base = smoothed_bridge(False, 101.0, 99.5, 100.0, 2.0, 0.04)
    up = smoothed_bridge(False, 101.0, 99.5, 100.0, 2.0, 0.04 + h)
    dn = smoothed_bridge(False, 101.0, 99.5, 100.0, 2.0, 0.04 - h)
    fd = (math.log(up[4]) - math.log(dn[4])) / (2 * h)
    assert abs((base[10] + base[11]) - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("case", PD["jump_cases"], ids=lambda c: c["id"])
def test_dividend_jump_cases(case):
    out = (C.c_double * 4)()
    oracle().oracle_jump_hit(int(case["direction"] == "up"), _d(case["pre_jump_spot"]), _d(case["post_jump_spot"]),
                             _d(case["barrier"]), _d(case["half_width"]), out)
    assert abs(out[0] - _d(case["expected"]["signed_score"])) <= 1e-14
    assert abs(out[1] - _d(case["expected"]["hit_weight"])) <= 1e-15
    assert abs(out[1] + out[2] - 1.0) <= 1e-15


def test_affine_dividend_event():
    c = G["affine_dividend"]
    i, e = c["inputs"], c["expected"]
    out = (C.c_double * 4)()
    oracle().oracle_dividend_event(_d(i["spot0"]), _d(i["a_before"]), _d(i["b_before"]), _d(i["fixed_cash"]),
                                   _d(i["proportional"]), _d(i["f"]), out)
    for k, g in zip(["alpha", "a_after", "b_after", "post_event_spot"], out):
        assert abs(g - _d(e[k])) <= 1e-14 * max(1.0, abs(_d(e[k]))), k


# ----------------------------------------------------------- replay pins --
def _bits(b):
    return struct.unpack(">d", bytes.fromhex(b))[0]


VALIDATION = {"delta": (abi.PE_C_BUMP_DELTA, abi.PE_C_DELTA_DIFFERENCE),
              "vega": (abi.PE_C_BUMP_VEGA, abi.PE_C_VEGA_DIFFERENCE),
              "gamma": (abi.PE_C_BUMP_GAMMA, abi.PE_C_GAMMA_DIFFERENCE)}
SKIP = {"pseudo_mc_delta_gamma_vega_vegakt", "rqmc_delta_gamma_vega_vegakt"}
VAL_CASES = [c for c in G["replay"] if c.get("risk_validation") and c["name"] not in SKIP]


@pytest.mark.parametrize("case", VAL_CASES, ids=[c["name"] for c in VAL_CASES])
def test_replay_bump_and_revalue_validation(case):
    res = oracle_evaluate(request.Request.from_v3(case["request"]), 4, case["reduction_block_size"])
    for key, (bump, diff) in VALIDATION.items():
        rv = case["risk_validation"].get(key)
        if not rv:
            continue
        want = _bits(rv["bump_and_revalue"]["value_bits"])
        # bump-and-revalue divides price differences by the bump: an ulp of
        # the price becomes ~1e-15 relative of the result, hence 1e-13.
        assert abs(res.estimates[bump].value - want) <= 1e-13 * abs(want), key
        assert abs(res.estimates[diff].value - _bits(rv["bump_minus_primary"]["value_bits"])) <= 1e-13 * abs(want)
        se = _bits(rv["bump_and_revalue"]["standard_error_bits"])
        assert abs(res.estimates[bump].standard_error - se) <= 1e-10 * se


def test_replay_continuous_barrier_diagnostics():
    case = next(c for c in G["replay"] if c["name"] == "continuous_barrier_smoothed_full_risk")
    res = oracle_evaluate(request.Request.from_v3(case["request"]), 4, case["reduction_block_size"])
    bb = case["barrier_bridge"]
    pairs = [("endpoint_hit_fraction_bits", abi.PE_C_BARRIER_ENDPOINT_HIT),
             ("dividend_jump_hit_fraction_bits", abi.PE_C_BARRIER_DIVIDEND_JUMP_HIT),
             ("mean_conditional_bridge_hit_weight_bits", abi.PE_C_BARRIER_BRIDGE_HIT_WEIGHT),
             ("mean_interval_count_bits", abi.PE_C_BARRIER_INTERVAL_COUNT),
             ("mean_finite_correction_count_bits", abi.PE_C_BARRIER_FINITE_CORRECTION_COUNT),
             ("mean_zero_variance_count_bits", abi.PE_C_BARRIER_ZERO_VARIANCE_COUNT),
             ("mean_survival_underflow_count_bits", abi.PE_C_BARRIER_SURVIVAL_UNDERFLOW_COUNT),
             ("mean_certain_survival_count_bits", abi.PE_C_BARRIER_CERTAIN_SURVIVAL_COUNT)]
    for key, comp in pairs:
        want = _bits(bb[key])
        assert abs(res.estimates[comp].value - want) <= 1e-13 * max(1.0, abs(want)), key


# ---------------------------------------------------------- closed forms --
R, Q, SIGMA, S0 = 0.03, 0.01, 0.25, 100.0


def _market(req, dividends=()):
    return req.set_market(S0, [0.0, 1.0], [1.0, math.exp(-R)], [0.0, 1.0], [1.0, math.exp(-Q)], dividends)


def reiner_rubinstein_up_out_call(s, k, b, r, q, v, t):
    """Continuously monitored up-and-out call, no rebate (Reiner-Rubinstein 1991; Haug, 'Option Pricing
    Formulas', 2nd ed., 4.17.1 with eta=-1, phi=1, K < H)."""
    mu = (r - q - 0.5 * v * v) / (v * v)
    sq = v * math.sqrt(t)
    x1 = math.log(s / k) / sq + (1 + mu) * sq
    x2 = math.log(s / b) / sq + (1 + mu) * sq
    y1 = math.log(b * b / (s * k)) / sq + (1 + mu) * sq
    y2 = math.log(b / s) / sq + (1 + mu) * sq
    eta, phi = -1.0, 1.0
    A = phi * s * math.exp(-q * t) * norm.cdf(phi * x1) - phi * k * math.exp(-r * t) * norm.cdf(phi * x1 - phi * sq)
    Bt = phi * s * math.exp(-q * t) * norm.cdf(phi * x2) - phi * k * math.exp(-r * t) * norm.cdf(phi * x2 - phi * sq)
    Ct = (phi * s * math.exp(-q * t) * (b / s) ** (2 * (mu + 1)) * norm.cdf(eta * y1)
          - phi * k * math.exp(-r * t) * (b / s) ** (2 * mu) * norm.cdf(eta * y1 - eta * sq))
    Dt = (phi * s * math.exp(-q * t) * (b / s) ** (2 * (mu + 1)) * norm.cdf(eta * y2)
          - phi * k * math.exp(-r * t) * (b / s) ** (2 * mu) * norm.cdf(eta * y2 - eta * sq))
    return A - Bt + Ct - Dt


def test_exact_continuous_barrier_matches_reiner_rubinstein():
    days = [73, 146, 219, 292, 365]
    req = _market(request.Request())
    req.set_black_scholes(SIGMA)
    req.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, 365, 100.0, 1.0, 365, barrier=130.0, continuous=1,
                    dates=days)
    req.set_pseudo(0x5EED, 1 << 15, antithetic=True)
    req.set_risk()
    res = oracle_evaluate(req, 8)
    closed = reiner_rubinstein_up_out_call(S0, 100.0, 130.0, R, Q, SIGMA, 1.0)
    est = res.estimates[0]
    assert abs(est.value - closed) <= 4.0 * est.standard_error, (est.value, closed, est.standard_error)
    # the diagnostics count one bridge interval per monitoring date
    assert res.estimates[abi.PE_C_BARRIER_INTERVAL_COUNT].value <= 5.0


def test_discrete_monitoring_overprices_the_continuous_barrier():
    days = [73, 146, 219, 292, 365]
    vals = []
    for cont in (0, 1):
        req = _market(request.Request())
        req.set_black_scholes(SIGMA)
        req.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, 365, 100.0, 1.0, 365, barrier=130.0, continuous=cont,
                        dates=days)
        req.set_pseudo(11, 1 << 13, antithetic=True)
        vals.append(oracle_evaluate(req, 8).estimates[0].value)
    assert vals[0] > vals[1] + 0.3


DIVS = [(0.25, 1.0, 0.01), (0.5, 2.0, 0.0), (0.75, 1.0, 0.02)]


def escrowed_forward(t):
    """Textbook forward with cash + proportional dividends:
    S0 e^{(r-q)t} prod(1-beta) - sum_i D_i e^{(r-q)(t-t_i)} prod_{j>i}(1-beta_j)."""
    f = S0 * math.exp((R - Q) * t)
    for (ti, _, beta) in DIVS:
        if ti <= t:
            f *= 1.0 - beta
    for i, (ti, cash, _) in enumerate(DIVS):
        if ti <= t:
            tail = 1.0
            for (tj, _, bj) in DIVS[i + 1:]:
                if tj <= t:
                    tail *= 1.0 - bj
            f -= cash * math.exp((R - Q) * (t - ti)) * tail
    return f


def test_escrowed_european_matches_black76_on_the_dividend_forward():
    req = _market(request.Request(), DIVS)
    req.set_black_scholes(SIGMA)
    req.set_product(abi.PE_PRODUCT_EUROPEAN, abi.PE_CALL, 365, 100.0, 1.0, 365)
    req.set_pseudo(3, 1 << 16, antithetic=True)
    req.set_risk(delta=True, vega=True)
    res = oracle_evaluate(req, 8)
    fwd, k, t, df = escrowed_forward(1.0), 100.0, 1.0, math.exp(-R)
    d1 = (math.log(fwd / k) + 0.5 * SIGMA * SIGMA * t) / (SIGMA * math.sqrt(t))
    d2 = d1 - SIGMA * math.sqrt(t)
    closed = df * (fwd * norm.cdf(d1) - k * norm.cdf(d2))
    est = res.estimates[0]
    assert abs(est.value - closed) <= 4.0 * est.standard_error
    # Delta holds every cash amount fixed (at_spot): dS_T/dS0 = prod(1-beta) F_T / S0.
    scale = (1 - 0.01) * (1 - 0.02)
    closed_delta = df * scale * math.exp((R - Q) * t) * norm.cdf(d1)
    d = res.estimates[1]
    assert abs(d.value - closed_delta) <= 4.0 * d.standard_error
    vega = df * fwd * norm.pdf(d1) * math.sqrt(t)
    v = res.estimates[2]
    assert abs(v.value - vega) <= 4.0 * v.standard_error


def test_escrowed_asian_mean_is_the_average_forward():
    days = [30, 91, 150, 200, 240, 300, 365]
    w = [1.0 / len(days)] * len(days)
    req = _market(request.Request(), DIVS)
    req.set_black_scholes(SIGMA)
    req.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, 365, 0.0, 1.0, 365, dates=days, weights=w,
                    fixings=[None] * len(days))
    req.set_pseudo(5, 1 << 15, antithetic=True)
    res = oracle_evaluate(req, 8)
    expected = math.exp(-R) * sum(wi * escrowed_forward(d / 365.0) for wi, d in zip(w, days))
    est = res.estimates[0]
    assert abs(est.value - expected) <= 4.0 * est.standard_error


def test_coordinates_reconstruct_the_dividend_forward():
    req = _market(request.Request(), DIVS)
    req.set_black_scholes(SIGMA)
    days = [30, 91, 183, 274, 365]
    req.set_product(abi.PE_PRODUCT_LOOKBACK, abi.PE_CALL, 365, 100.0, 1.0, 365, dates=days)
    req.set_pseudo(1, 16)
    rows = (C.c_double * (8 * 16))()
    reserve = C.c_double()
    n = oracle().oracle_plan_coordinates(req.pointer(), rows, 16, C.byref(reserve))
    assert n == len(days)
    for i in range(n):
        t, a, b = rows[8 * i], rows[8 * i + 1], rows[8 * i + 2]
        plain = S0 * math.exp((R - Q) * t)
        assert abs(a * S0 + b * plain - escrowed_forward(t)) <= 1e-12 * S0


def test_zero_vol_barrier_is_hit_before_the_dividend_jump():
    # Up-and-out: the pre-dividend spot on the ex-date (day 182) is above the
    # barrier, the post-dividend spot below it.  Monitoring both sides knocks out.
    divs = [(182 / 365.0, 0.0, 0.05)]
    req = request.Request().set_market(S0, [0.0, 1.0], [1.0, math.exp(-0.1)], [0.0, 1.0], [1.0, 1.0], divs)
    req.set_black_scholes(0.0)
    fwd = S0 * math.exp(0.1 * 182 / 365.0)
    barrier = fwd - 0.5 * (fwd - 0.95 * fwd)  # between post and pre spot
    req.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, 365, 50.0, 1.0, 365, barrier=barrier, dates=[182, 365],
                    rebate=3.0)
    req.set_pseudo(1, 4)
    res = oracle_evaluate(req, 1)
    assert res.estimates[0].value == pytest.approx(3.0 * math.exp(-0.1), rel=1e-14)
    # Moving the barrier above the pre-dividend spot keeps the option alive.
    req.c.product.barrier = fwd * 1.01
    res = oracle_evaluate(req, 1)
    assert res.estimates[0].value > 3.0
