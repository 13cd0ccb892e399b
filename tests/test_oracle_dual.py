"""Andersen-Broadie primal-dual bounds in the oracle (the reference's
AndersenBroadiePlan, crates/pricing/src/engine/risk/dual.rs).  Pins: the
reference's own acceptance properties (crates/pricing/tests/american_dual.rs)
-- one exercise date gives a zero gap and the plain Monte Carlo value, a
deterministic grid gives lower = upper = the best discounted payoff, the lower
bound is the frozen policy's LSM price -- and the bracket around a CRR tree
value."""
import math

import pytest

from _support import oracle_lsm, oracle_lsm_dual
from pricing_amd import abi, request
from test_oracle_lsm import crr_bermudan_put

DATES = [91, 181, 273, 365]   # 2026-12-04 .. 2027-09-04 from 2026-09-04


def american(vol=0.2, strike=100.0, dates=DATES, units=256, training_units=2048, antithetic=True, seed=91,
             method=abi.PE_LSM_POLY_CPQR, side=abi.PE_PUT, rate=0.05):
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-rate)], [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(vol)
    req.set_product(abi.PE_PRODUCT_AMERICAN, side, dates[-1], strike, dates=dates)
    req.set_pseudo(seed, units, antithetic=antithetic)
    req.set_lsm((abi.PE_ENGINE_PSEUDO, seed + 1, training_units, 0), antithetic=antithetic, method=method)
    req.set_risk()
    return req


def test_single_exercise_date_has_zero_gap_and_the_plain_value():
    req = american(dates=[365])
    lower, upper, gap = oracle_lsm_dual(req, 64, 33, 0x1234)
    price = oracle_lsm(req, 64)[0].estimates[0]
    assert abs(lower.value - price.value) < 1e-13
    assert lower.value == upper.value and gap.value == 0.0 and gap.standard_error == 0.0


@pytest.mark.parametrize("strike", [120.0, 80.0])
def test_deterministic_grid_gives_the_best_discounted_payoff(strike):
    req = american(vol=0.0, strike=strike, side=abi.PE_PUT if strike > 100 else abi.PE_CALL, rate=0.05)
    lower, upper, gap = oracle_lsm_dual(req, 64, 33, 0x1234)
    best = 0.0
    for day in DATES:
        t = day / 365.0
        spot = 100.0 * math.exp(0.05 * t)
        payoff = max(strike - spot, 0.0) if strike > 100 else max(spot - strike, 0.0)
        best = max(best, payoff * math.exp(-0.05 * t))
    assert abs(upper.value - best) < 1e-11 and abs(lower.value - best) < 1e-11
    assert upper.standard_error == 0.0


def test_lower_bound_is_the_frozen_policy_price():
    req = american(units=256)
    lower, upper, gap = oracle_lsm_dual(req, 128, 65, 0x1234, block=64)
    price = oracle_lsm(req, 64)[0].estimates[0]
    assert abs(lower.value - price.value) < 1e-12 * abs(price.value)
    assert upper.value >= lower.value and gap.value >= 0.0


def test_bracket_contains_the_tree_value():
    req = american(units=2048, training_units=8192, method=abi.PE_LSM_POLY_LAGUERRE)
    lower, upper, gap = oracle_lsm_dual(req, 256, 129, 0x5EED, block=256, workers=8)
    tree = crr_bermudan_put(100.0, 100.0, 0.05, 0.2, 1.0, [d / 365.0 for d in DATES], steps=4000)
    assert lower.value - 4.0 * lower.standard_error <= tree + 0.01
    assert upper.value + 4.0 * upper.standard_error >= tree - 0.01
    assert gap.value < 0.1, gap.value
