"""GPU checks of the allocation-free entry point pe_plan_evaluate_noalloc:
through the C ABI it returns what pe_plan_evaluate returns for the same plan
(the same kernel on the same items: bit for bit), on two calls in a row with
the second one smaller, in both device modes; a plan too large for the fixed
workspace and a plan kind outside its scope are rejected before any launch.
"""
import math

import pytest

from pricing_amd import abi, engine, request, workloads

pytestmark = pytest.mark.gpu

MONTHLY = [30, 61, 91, 122, 152, 182, 213, 243, 274, 304, 335, 365]


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def _continuous_barrier(units):
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-0.03)], [0.0, 1.0], [1.0, 1.0],
                                       [(0.25, 1.0, 0.01), (0.5, 2.0, 0.0)])
    req.set_black_scholes(0.25)
    req.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, 365, 100.0, barrier=130.0, continuous=1, dates=MONTHLY,
                    rebate=2.0)
    req.set_pseudo(0x0123456789ABCDEF, units, antithetic=True)
    req.set_risk(delta=True, vega=True, gamma_bump=0.01, smoothing_half_width=2.0)
    return req


def _asian_rqmc(points, scrambles):
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-0.03)], [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(0.2)
    req.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, 365, 100.0, dates=MONTHLY, weights=[1 / 12] * 12)
    req.set_rqmc(0xFEDCBA9876543210, points, scrambles, antithetic=True)
    req.set_risk(delta=True, vega=True)
    return req


def _same(a, b):
    for c in range(abi.COMPONENTS):
        assert a.estimates[c].value == b.estimates[c].value, (c, a.estimates[c].value, b.estimates[c].value)
        assert a.estimates[c].standard_error == b.estimates[c].standard_error, c
    assert a.sampling_variance == b.sampling_variance
    assert a.evaluated_paths == b.evaluated_paths and a.path_steps == b.path_steps


CASES = [
    ("european risk 2^16 then 2^10", lambda: workloads.european_bs(units=1 << 16),
     lambda: workloads.european_bs(units=1 << 10)),
    ("continuous barrier 2^15 then 2^11", lambda: _continuous_barrier(1 << 15), lambda: _continuous_barrier(1 << 11)),
    ("asian rqmc 2^12x8 then 2^9x2", lambda: _asian_rqmc(1 << 12, 8), lambda: _asian_rqmc(1 << 9, 2)),
    ("graph asian 2^16 then 2^12", lambda: workloads.graph_asian(units=1 << 16),
     lambda: workloads.graph_asian(units=1 << 12)),
]


@pytest.mark.parametrize("backend", ["fp64", "mixed"])
@pytest.mark.parametrize("name,big,small", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_evaluate_on_two_calls(name, big, small, backend):
    print("case", name, backend, flush=True)
    for make in (big, small):
        plan = engine.PricingPlan.compile(make())
        first = plan.evaluate_noalloc(backend)
        reference = plan.evaluate(backend)
        _same(first, reference)


def test_noalloc_rejects_a_plan_beyond_the_workspace():
    # 2^27 units with risk: 32768 work items x 10 statistics x 40 B > 8 MiB
    plan = engine.PricingPlan.compile(workloads.european_bs(units=1 << 27))
    with pytest.raises(engine.EngineError) as err:
        plan.evaluate_noalloc("fp64")
    assert err.value.status == abi.PE_ERR_SIZE_LIMIT
    # the workspace is still usable afterwards
    small = engine.PricingPlan.compile(workloads.european_bs(units=1 << 12))
    _same(small.evaluate_noalloc("fp64"), small.evaluate("fp64"))


def test_noalloc_rejects_other_plan_kinds():
    model = workloads.lifted_heston_model(engine.lifted_from_rough(0.1, 20, 2.5))
    carrier = workloads.rough_smile(points=256, scrambles=0, strikes=2, maturities=(0.25,))[0]
    plan = engine.RoughPlan.compile(carrier, model, 2.0 / 256)
    with pytest.raises(engine.EngineError) as err:
        plan.evaluate_noalloc("fp64")
    assert err.value.status == abi.PE_ERR_UNSUPPORTED
