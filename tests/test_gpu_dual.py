"""Andersen-Broadie bounds on the device (the LSM training launch, the inner
rollout and outer kernels) against the oracle: FP64 lower bound, upper bound
and duality gap with their standard errors within 1e-12 relative, and the
reference's acceptance properties (crates/pricing/tests/american_dual.rs)."""
import pytest

from _support import oracle_lsm, oracle_lsm_dual
from pricing_amd import abi, engine
from test_oracle_dual import DATES, american

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def device(req, n2, n3, seed):
    return engine.evaluate_lsm_dual(engine.PricingPlan.compile(req), n2, n3, seed)


CASES = [
    ("cpqr-antithetic", dict(units=256), 128, 65),
    ("laguerre", dict(units=512, method=abi.PE_LSM_POLY_LAGUERRE, antithetic=False), 96, 33),
    ("call-monthly", dict(units=256, side=abi.PE_CALL, dates=[30 * k for k in range(1, 13)]), 64, 64),
]


@pytest.mark.parametrize("name,kw,n2,n3", CASES, ids=[c[0] for c in CASES])
def test_dual_fp64_matches_oracle(name, kw, n2, n3):
    req = american(**kw)
    print("case", name, flush=True)
    dev = device(req, n2, n3, 0x1234)
    ref = oracle_lsm_dual(req, n2, n3, 0x1234, block=256, workers=8)
    for got, exp, what in ((dev.lower_bound, ref[0], "lower"), (dev.upper_bound, ref[1], "upper"),
                           (dev.duality_gap, ref[2], "gap")):
        assert abs(got.value - exp.value) <= 1e-12 * max(abs(exp.value), 1e-300) + 1e-15, (what, got.value, exp.value)
        assert abs(got.standard_error - exp.standard_error) <= 1e-10 * exp.standard_error + 1e-15, what
    assert dev.outer_trajectories == req.c.engine.units * (2 if req.c.engine.antithetic else 1)


def test_single_exercise_date_has_zero_gap_and_the_plain_value():
    req = american(dates=[365])
    dev = device(req, 64, 33, 0x1234)
    price = engine.evaluate_lsm([engine.PricingPlan.compile(req)], "fp64")[0][0].estimates[0]
    assert abs(dev.lower_bound.value - price.value) < 1e-13
    assert dev.duality_gap.value == 0.0 and dev.upper_bound.value == dev.lower_bound.value


def test_zero_volatility_bounds_coincide():
    req = american(vol=0.0, strike=120.0)
    dev = device(req, 64, 33, 0x1234)
    ref = oracle_lsm_dual(req, 64, 33, 0x1234)
    assert abs(dev.upper_bound.value - ref[1].value) < 1e-12 and dev.upper_bound.standard_error == 0.0
    assert abs(dev.upper_bound.value - dev.lower_bound.value) < 1e-12


def test_lower_bound_is_the_device_lsm_price():
    req = american(units=1024, training_units=4096, method=abi.PE_LSM_POLY_LAGUERRE)
    dev = device(req, 128, 64, 0x77)
    price = engine.evaluate_lsm([engine.PricingPlan.compile(req)], "fp64")[0][0].estimates[0]
    assert abs(dev.lower_bound.value - price.value) <= 1e-12 * abs(price.value)
    assert dev.upper_bound.value >= dev.lower_bound.value


def test_unsupported_requests_are_rejected():
    req = american()
    req.set_risk(delta=True)
    with pytest.raises(Exception):
        device(req, 64, 33, 1)
    with pytest.raises(Exception):
        engine.evaluate_lsm_dual(engine.PricingPlan.compile(american()), 64, 33, 1, backend="mixed")
