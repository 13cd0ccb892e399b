"""Pins the oracle's market layer: SSVI derivatives and Dupire local variance
against the reference's Decimal fixtures, LV time grids and node helpers
against their definitions, and the BS path against the closed form."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.stats import norm

from _support import darr, golden, oracle, oracle_evaluate
from pricing_amd import abi, request


def ssvi(theta_knots, theta_values, slope, rho, eta, gamma):
    s = abi.pe_ssvi_surface()
    s.theta_count = len(theta_knots)
    s._t = darr(theta_knots)
    s._v = darr(theta_values)
    s.theta_times = C.cast(s._t, C.POINTER(C.c_double))
    s.theta_values = C.cast(s._v, C.POINTER(C.c_double))
    s.theta_terminal_slope = slope
    s.rho, s.eta, s.gamma = rho, eta, gamma
    return s


def test_standard_ssvi_power_law_decimal_case():
    # fixtures/local-vol/reference-cases-v0.1.json standard_ssvi/power_regular:
    # theta = 0.04, theta_t = 0.03 at the query time.  A two-knot PCHIP through
    # (1, 0.04) and (2, 0.07) has derivative 0.03 at its first knot.
    case = golden()["standard_ssvi"][0]
    i, e = case["inputs"], case["expected"]
    s = ssvi([1.0, 2.0], [0.04, 0.07], 0.03, float(i["rho"]), float(i["eta"]), float(i["gamma"]))
    out = (C.c_double * 4)()
    assert oracle().oracle_ssvi_derivatives(C.byref(s), 1.0, float(i["k"]), out) == 0
    w, wk, wkk, wt = out
    for got, key in ((w, "w"), (wk, "w_k"), (wkk, "w_kk"), (wt, "w_t")):
        assert abs(got - float(e[key])) <= 1e-15 * max(1.0, abs(float(e[key]))), key
    k = float(i["k"])
    u = 1.0 - k * wk / (2.0 * w)
    g = u * u - (wk * wk / 4.0) * (1.0 / w + 0.25) + wkk / 2.0
    assert abs(g - float(e["density_factor"])) < 1e-15


def test_local_variance_grid_equals_wt_over_density_factor():
    case = golden()["standard_ssvi"][0]
    i, e = case["inputs"], case["expected"]
    s = ssvi([1.0, 2.0], [0.04, 0.07], 0.03, float(i["rho"]), float(i["eta"]), float(i["gamma"]))
    t = [0.0, 1.0]
    x = [float(i["k"]), 0.2]
    out = (C.c_double * 4)()
    rep = C.c_uint32()
    assert oracle().oracle_local_variance_from_ssvi(C.byref(s), 2, darr(t), 2, darr(x), 1e-8, 4.0, out,
                                                   C.byref(rep)) == 0
    expected = float(e["w_t"]) / float(e["density_factor"])
    # t = 0 row evaluates the surface at the first positive node (local_variance.rs:166).
    assert abs(out[0] - expected) < 1e-15
    assert abs(out[2] - expected) < 1e-15
    assert rep.value == 0


def test_constant_vol_dupire_identity():
    # fixtures/local-vol/reference-cases-v0.1.json constant_vol_dupire: with a
    # flat smile (rho = 0, phi -> 0 is not representable), a pure-theta SSVI at
    # k = 0 reduces to w = theta, local variance = dtheta/dt.
    s = ssvi([0.5, 2.0], [0.02, 0.08], 0.04, 0.0, 1e-9, 0.4)
    out = (C.c_double * 1)()
    rep = C.c_uint32()
    assert oracle().oracle_local_variance_from_ssvi(C.byref(s), 2, darr([0.0, 1.25]), 1, darr([0.0]), 1e-8, 4.0,
                                                   out, C.byref(rep)) != 0  # one x node is invalid
    out = (C.c_double * 4)()
    assert oracle().oracle_local_variance_from_ssvi(C.byref(s), 2, darr([0.0, 1.25]), 2, darr([0.0, 0.3]), 1e-8,
                                                   4.0, out, C.byref(rep)) == 0
    case = golden()["constant_vol_dupire"]
    assert abs(out[2] - float(case["expected"]["local_variance"])) < 1e-9


def test_repairs_floor_and_cap():
    s = ssvi([1.0, 2.0], [0.04, 0.07], 0.03, -0.3, 0.5, 0.5)
    out = (C.c_double * 2)()
    rep = C.c_uint32()
    assert oracle().oracle_local_variance_from_ssvi(C.byref(s), 2, darr([0.0, 1.0]), 1 + 1, darr([-0.1, 0.1]),
                                                   0.5, 4.0, (C.c_double * 4)(), C.byref(rep)) == 0
    assert rep.value == 4  # every raw value (~0.03) is below the 0.5 floor


def time_grid(events, max_step):
    nodes = (C.c_double * 4096)()
    n = C.c_uint32()
    assert oracle().oracle_lv_time_grid(darr(events), len(events), max_step, nodes, 4096, C.byref(n)) == 0
    return list(nodes[: n.value])


def test_lv_time_grid_union_and_substeps():
    # local_vol.rs:166-237: union of events, ceil(dt/max) equal substeps.
    g = time_grid([1.0, 0.5, 0.5, 0.2], 0.2)
    assert g[0] == 0.0 and g[-1] == 1.0
    assert 0.2 in g and 0.5 in g
    # [0,0.2]: 1 step; [0.2,0.5]: 2 steps; [0.5,1.0]: 3 steps
    assert len(g) == 1 + 1 + 2 + 3
    assert max(b - a for a, b in zip(g, g[1:])) <= 0.2 + 1e-15
    with pytest.raises(AssertionError):
        time_grid([-0.0, 1.0], 0.5)


def test_sinh_nodes_endpoints_and_monotone():
    out = (C.c_double * 64)()
    oracle().oracle_sinh_nodes(-1.5, 1.0, 32, 31, 3.0, 3.0, out)
    v = list(out)
    assert v[0] == -1.5 and v[32] == 0.0 and abs(v[63] - 1.0) < 1e-15
    assert all(b > a for a, b in zip(v, v[1:]))
    # shape 0 gives a uniform grid
    oracle().oracle_sinh_nodes(-1.0, 1.0, 4, 4, 0.0, 0.0, out)
    assert list(out[:9]) == [-1.0, -0.75, -0.5, -0.25, 0.0, 0.25, 0.5, 0.75, 1.0]


def bs_request(units, seed=0x0123456789ABCDEF, sigma=0.2, strike=100.0, r=0.05, q=0.0, side=abi.PE_CALL):
    req = request.Request()
    req.set_market(100.0, [0.0, 1.0], [1.0, math.exp(-r)], [0.0, 1.0], [1.0, math.exp(-q)])
    req.set_black_scholes(sigma)
    req.set_product(abi.PE_PRODUCT_EUROPEAN, side, 365, strike)
    req.set_pseudo(seed, units)
    req.set_risk(delta=True, vega=True)
    return req


def bs_closed_form(s, k, r, q, sigma, t, call=True):
    f = s * math.exp((r - q) * t)
    sd = sigma * math.sqrt(t)
    d1 = (math.log(f / k) + 0.5 * sd * sd) / sd
    d2 = d1 - sd
    df = math.exp(-r * t)
    if call:
        return df * (f * norm.cdf(d1) - k * norm.cdf(d2)), math.exp(-q * t) * norm.cdf(d1), s * math.exp(
            -q * t) * norm.pdf(d1) * math.sqrt(t)
    return df * (k * norm.cdf(-d2) - f * norm.cdf(-d1)), -math.exp(-q * t) * norm.cdf(-d1), s * math.exp(
        -q * t) * norm.pdf(d1) * math.sqrt(t)


@pytest.mark.parametrize("side", [abi.PE_CALL, abi.PE_PUT])
def test_european_bs_price_delta_vega_within_se_of_closed_form(side):
    # The small-sample version of the workload S0=K=100, r=5%, sigma=20%, T=1.
    res = oracle_evaluate(bs_request(1 << 16, side=side), workers=8)
    price, delta, vega = bs_closed_form(100, 100, 0.05, 0.0, 0.2, 1.0, side == abi.PE_CALL)
    for idx, ref in ((0, price), (1, delta), (2, vega)):
        est = res.estimates[idx]
        assert abs(est.value - ref) < 5 * est.standard_error, (idx, est.value, ref, est.standard_error)
    # bump validation agrees with the pathwise estimate
    assert abs(res.estimates[5].value) < 5 * res.estimates[5].standard_error + 1e-6
