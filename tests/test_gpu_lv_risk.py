"""Local-vol Delta, Gamma and Vega on the device (lv_risk_kernel) against the
reference's replay values and the oracle.

Delta and Gamma re-simulate the shocks from spot -/+ the validation bump; the
FP64 mode follows the reference's recursion for the bumped runtimes, so their
bar is the FP64 one (1e-12 of the rounding scale price / h, h^2 for Gamma).
Vega is the derivative along every local-variance node's volatility: the
reference sums adjoints of a reverse sweep, the device carries the tangent
forward with the path, so the two agree to rounding (1e-12 relative).
Mixed mode: within 0.1 SE of the FP64 mode.
"""
import math

import pytest

from _support import golden, oracle_evaluate
from pricing_amd import abi, engine, request

pytestmark = pytest.mark.gpu

REL = 1e-12
RISK_CASES = [c for c in golden()["replay"] if c["name"] in (
    "pseudo_mc_delta_gamma_vega_vegakt", "rqmc_delta_gamma_vega_vegakt")]


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def bump_of(req):
    spot = req.c.market.spot
    if req.c.risk.gamma:
        return spot * req.c.risk.gamma_bump if req.c.risk.gamma_relative else req.c.risk.gamma_bump
    return spot * 1e-4


def check(gpu_value, ref_value, scale, what):
    assert abs(gpu_value - ref_value) <= REL * scale + 1e-15, (what, gpu_value, ref_value, scale)


def compare_risk(gpu, ref, req, se_rel=1e-10):
    price = abs(ref.estimates[0].value)
    h = bump_of(req)
    scales = {0: price, 1: abs(ref.estimates[1].value) + price / h, 2: abs(ref.estimates[2].value),
              3: abs(ref.estimates[3].value) + price / (h * h)}
    for c, scale in scales.items():
        check(gpu.estimates[c].value, ref.estimates[c].value, scale, abi.COMPONENT_NAMES[c])
        r = ref.estimates[c].standard_error
        g = gpu.estimates[c].standard_error
        se_scale = r * scale / max(abs(ref.estimates[c].value), 1e-300)
        assert abs(g - r) <= max(se_rel * max(r, se_scale), 1e-15), (abi.COMPONENT_NAMES[c], g, r)


@pytest.mark.parametrize("case", RISK_CASES, ids=[c["name"] for c in RISK_CASES])
def test_replay_local_vol_risk_on_device(case):
    req = request.Request.from_v3(case["request"])
    res = engine.PricingPlan.compile(req).evaluate("fp64")
    price = case["value"]["value"]
    assert abs(res.estimates[0].value - price) <= REL * abs(price)
    h = bump_of(req)
    for key, idx, scale in (("delta", 1, abs(price) / h), ("vega", 2, 0.0), ("gamma", 3, abs(price) / h / h)):
        e = case["risks"][key]
        check(res.estimates[idx].value, e["value"], abs(e["value"]) + scale, key)
        assert abs(res.estimates[idx].standard_error - e["standard_error"]) <= 1e-9 * e["standard_error"], key
    assert res.estimates[abi.PE_C_BUMP_DELTA].value == res.estimates[1].value
    assert res.estimates[abi.PE_C_BUMP_GAMMA].value == res.estimates[3].value


def lv_request(kind, engine_kind="pseudo", bridge=False, units=1 << 12, scrambles=4, steps=24, gamma=0.01,
               delta=True, vega=True):
    """Small local-vol trades on the replay fixtures' 3 x 5 grid, T = 1y."""
    req = request.Request()
    r = 0.03
    req.set_market(100.0, [0.0, 1.0], [1.0, math.exp(-r)], [0.0, 1.0], [1.0, 1.0])
    ts = [0.0, 0.4986301369863014, 1.0]
    xs = [-0.3, -0.1, 0.0, 0.2, 0.4]
    vs = [0.048, 0.042, 0.039, 0.041, 0.047, 0.051, 0.044, 0.04, 0.043, 0.05, 0.056, 0.049, 0.045, 0.047, 0.053]
    req.set_local_vol(ts, xs, vs, 1e-8, 4.0)
    dates = [round(365 * (k + 1) / steps) for k in range(steps)]
    if kind == "european_put":
        req.set_product(abi.PE_PRODUCT_EUROPEAN, abi.PE_PUT, 365, 105.0)
    elif kind == "barrier_smooth":
        req.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, 365, 100.0, barrier=125.0, dates=dates)
    elif kind == "barrier_down_in":
        req.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_PUT, 365, 100.0, barrier=85.0, direction_down=1, knock_in=1,
                        dates=dates)
    elif kind == "asian":
        req.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, 365, 100.0, dates=dates, weights=[1.0 / steps] * steps)
    elif kind == "lookback":
        req.set_product(abi.PE_PRODUCT_LOOKBACK, abi.PE_CALL, 365, 100.0, dates=dates)
    elif kind == "digital_smooth":
        req.set_product(abi.PE_PRODUCT_DIGITAL, abi.PE_CALL, 365, 100.0, payout=1.0)
    else:
        raise ValueError(kind)
    if engine_kind == "pseudo":
        req.set_pseudo(0xA5A5_1234, units, antithetic=True, brownian_bridge=bridge)
    else:
        req.set_rqmc(0x0BAD_CAFE, units, scrambles, antithetic=False, brownian_bridge=bridge)
    smooth = 1.0 if kind in ("barrier_smooth", "digital_smooth") else None
    req.set_risk(delta=delta, vega=vega, gamma_bump=gamma, smoothing_half_width=smooth)
    return req


CASES = [
    ("european_put", "pseudo", False),
    ("barrier_smooth", "rqmc", True),
    ("barrier_down_in", "pseudo", True),
    ("asian", "rqmc", False),
    ("lookback", "pseudo", False),
    ("digital_smooth", "rqmc", True),
]


@pytest.mark.parametrize("kind,engine_kind,bridge", CASES, ids=[c[0] for c in CASES])
def test_fp64_matches_oracle(kind, engine_kind, bridge):
    req = lv_request(kind, engine_kind, bridge)
    print("case", kind, engine_kind, bridge, flush=True)
    gpu = engine.PricingPlan.compile(req).evaluate("fp64")
    ref = oracle_evaluate(req, workers=8)
    compare_risk(gpu, ref, req)


def test_partial_requests_fill_only_what_is_asked():
    req = lv_request("european_put", delta=False, vega=True, gamma=None)
    gpu = engine.PricingPlan.compile(req).evaluate("fp64")
    ref = oracle_evaluate(req, workers=8)
    assert gpu.estimates[1].value == 0.0 and gpu.estimates[3].value == 0.0
    check(gpu.estimates[2].value, ref.estimates[2].value, abs(ref.estimates[2].value), "vega")
    req = lv_request("asian", delta=True, vega=False, gamma=None)
    gpu = engine.PricingPlan.compile(req).evaluate("fp64")
    ref = oracle_evaluate(req, workers=8)
    assert gpu.estimates[2].value == 0.0
    check(gpu.estimates[1].value, ref.estimates[1].value, abs(ref.estimates[1].value) +
          abs(ref.estimates[0].value) / bump_of(req), "delta")


@pytest.mark.parametrize("kind,engine_kind,bridge", CASES, ids=[c[0] for c in CASES])
def test_mixed_within_tenth_se(kind, engine_kind, bridge):
    req = lv_request(kind, engine_kind, bridge, units=1 << 14)
    plan = engine.PricingPlan.compile(req)
    f64 = plan.evaluate("fp64")
    mixed = plan.evaluate("mixed")
    for c in (0, 1, 2, 3):
        se = f64.estimates[c].standard_error
        assert abs(mixed.estimates[c].value - f64.estimates[c].value) <= 0.1 * se + 1e-12, (
            abi.COMPONENT_NAMES[c], mixed.estimates[c].value, f64.estimates[c].value, se)
