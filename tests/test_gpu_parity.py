"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle and
the reference's golden replay fixtures.

Bars: FP64 device mode within 1e-12 relative of the oracle on prices, Greeks
and standard errors; mixed mode within 0.1 standard errors of the FP64 mode.
"""
import math

import pytest

from _support import golden, oracle, oracle_evaluate
from pricing_amd import abi, engine, request, workloads

pytestmark = pytest.mark.gpu

REL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def close(a, b, rel=REL, abs_tol=1e-15):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + abs_tol


def conditioning(req, ref, c):
    """Absolute scale of the rounding a component inherits.  Prices and the
    pathwise Greeks are well conditioned (scale = |value|).  The bump-and-
    revalue validation components divide price differences by the bump h
    (h^2 for the second difference), so a price-level ulp becomes
    ulp * |price| / h: their bar is 1e-12 of that scale."""
    price = abs(ref.estimates[0].value) + 1e-300
    spot = req.c.market.spot
    spot_bump = spot * req.c.risk.gamma_bump if req.c.risk.gamma else spot * 1e-4
    vol = req.c.model.volatility
    vol_bump = min(1e-4, 0.5 * vol) if vol else 1.0
    scale = {4: price / spot_bump, 5: price / spot_bump, 6: price / vol_bump, 7: price / vol_bump,
             8: price / spot_bump**2, 9: price / spot_bump**2}
    return scale.get(c, 0.0) + abs(ref.estimates[c].value)


def compare(gpu, ref, components, rel=REL, req=None):
    for c in components:
        g, r = gpu.estimates[c], ref.estimates[c]
        scale = conditioning(req, ref, c) if req is not None else abs(r.value)
        assert abs(g.value - r.value) <= rel * scale + 1e-15, (abi.COMPONENT_NAMES[c], g.value, r.value, scale)
        se_scale = (scale / max(abs(r.value), 1e-300)) * r.standard_error if req is not None else r.standard_error
        assert abs(g.standard_error - r.standard_error) <= max(rel, 1e-10) * max(se_scale, r.standard_error) + 1e-15, (
            abi.COMPONENT_NAMES[c], g.standard_error, r.standard_error)


RISK = list(range(10))


def test_european_bs_fp64_matches_oracle():
    req = workloads.european_bs(units=1 << 14)
    gpu = engine.PricingPlan.compile(req).evaluate("fp64")
    ref = oracle_evaluate(req, workers=8)
    compare(gpu, ref, RISK, req=req)
    assert gpu.evaluated_paths == 1 << 14


def test_european_bs_mixed_within_tenth_se():
    req = workloads.european_bs(units=1 << 16)
    plan = engine.PricingPlan.compile(req)
    f64 = plan.evaluate("fp64")
    mixed = plan.evaluate("mixed")
    for c in (0, 1, 2):
        assert abs(mixed.estimates[c].value - f64.estimates[c].value) < 0.1 * f64.estimates[c].standard_error


GPU_REPLAY = [c for c in golden()["replay"] if c["name"] not in (
    "pseudo_mc_delta_gamma_vega_vegakt", "rqmc_delta_gamma_vega_vegakt")]


@pytest.mark.parametrize("case", GPU_REPLAY, ids=[c["name"] for c in GPU_REPLAY])
def test_replay_fixture_on_device(case):
    req = request.Request.from_v3(case["request"])
    res = engine.PricingPlan.compile(req).evaluate("fp64")
    exp = case["value"]
    assert close(res.estimates[0].value, exp["value"]), (res.estimates[0].value, exp["value"])
    assert close(res.estimates[0].standard_error, exp["standard_error"], 1e-10)
    for key, idx in (("delta", 1), ("vega", 2), ("gamma", 3)):
        if key in case["risks"]:
            e = case["risks"][key]
            assert close(res.estimates[idx].value, e["value"]), (key, res.estimates[idx].value, e["value"])
    # the evaluation reports the fixture's RQMC checksums and engine identity
    mc, d = case["monte_carlo"], res.diagnostics
    assert d.master_seed == int(mc["master_seed"]) and d.scramble_count == (mc["scramble_count"] or 0)
    if mc["direction_checksum"]:
        assert bytes(d.direction_checksum).hex() == mc["direction_checksum"]
        assert bytes(d.scramble_checksum).hex() == mc["scramble_checksum"]


def test_rqmc_points_not_multiple_of_wave():
    case = next(c for c in golden()["replay"] if c["name"] == "rqmc_full_risk")
    req = request.Request.from_v3(case["request"])
    req.c.engine.units = 1000
    compare(engine.PricingPlan.compile(req).evaluate("fp64"), oracle_evaluate(req, workers=8), RISK, req=req)


def _lv_requests(points, scrambles, steps=252):
    grid = workloads.lv_dupire_grid(steps_expiry(steps), build=engine.local_variance_from_ssvi)
    return workloads.lv_path_dependent(grid, points=points, scrambles=scrambles, steps=steps)


def steps_expiry(steps):
    return workloads.weekday_dates(steps)[-1] / 365.0


def test_dupire_grid_device_matches_oracle():
    import ctypes as C

    expiry = steps_expiry(252)
    gpu = workloads.lv_dupire_grid(expiry, build=engine.local_variance_from_ssvi)

    def oracle_build(tt, tv, slope, rho, eta, gamma, times, xs, floor, cap):
        s = abi.pe_ssvi_surface()
        a = (C.c_double * 2)(*tt)
        b = (C.c_double * 2)(*tv)
        s.theta_count, s.theta_times, s.theta_values = 2, C.cast(a, C.POINTER(C.c_double)), C.cast(
            b, C.POINTER(C.c_double))
        s.theta_terminal_slope, s.rho, s.eta, s.gamma = slope, rho, eta, gamma
        out = (C.c_double * (len(times) * len(xs)))()
        rep = C.c_uint32()
        assert oracle().oracle_local_variance_from_ssvi(C.byref(s), len(times), (C.c_double * len(times))(*times),
                                                       len(xs), (C.c_double * len(xs))(*xs), floor, cap, out,
                                                       C.byref(rep)) == 0
        return list(out), rep.value

    ref = workloads.lv_dupire_grid(expiry, build=oracle_build)
    assert gpu[5] == ref[5]
    for g, r in zip(gpu[2], ref[2]):
        assert close(g, r, 1e-13), (g, r)


@pytest.mark.parametrize("points,scrambles", [(1 << 12, 4), (1000, 2)])
def test_local_vol_rqmc_bridge_fp64_matches_oracle(points, scrambles):
    reqs = _lv_requests(points, scrambles)
    plans = [engine.PricingPlan.compile(r) for r in reqs]
    gpu = engine.evaluate_shared(plans, "fp64")
    import ctypes as C

    lib = oracle()
    arr = (C.POINTER(abi.pe_request) * 2)(*[r.pointer() for r in reqs])
    ref = (abi.pe_result * 2)()
    pol = request.exec_policy(8, 0)
    assert lib.oracle_evaluate_shared(arr, 2, C.byref(pol), ref) == 0, lib.oracle_last_error()
    for g, r in zip(gpu, ref):
        compare(g, r, [0])
    # single-plan evaluation equals the shared one
    single = plans[1].evaluate("fp64")
    assert single.estimates[0].value == gpu[1].estimates[0].value


def test_local_vol_mixed_within_tenth_se():
    reqs = _lv_requests(1 << 14, 8)
    plans = [engine.PricingPlan.compile(r) for r in reqs]
    f64 = engine.evaluate_shared(plans, "fp64")
    mixed = engine.evaluate_shared(plans, "mixed")
    for a, b in zip(f64, mixed):
        assert abs(a.estimates[0].value - b.estimates[0].value) < 0.1 * a.estimates[0].standard_error, (
            a.estimates[0].value, b.estimates[0].value, a.estimates[0].standard_error)


def test_local_vol_pseudo_bridge_and_step_order_match_oracle():
    case = next(c for c in golden()["replay"] if c["name"] == "pseudo_mc_price_only")
    for bridge in (0, 1):
        req = request.Request.from_v3(case["request"])
        req.c.engine.brownian_bridge = bridge
        req.c.engine.units = 3000
        compare(engine.PricingPlan.compile(req).evaluate("fp64"), oracle_evaluate(req, workers=8), [0])


def test_normal_mixed_device_accuracy():
    """The mixed kernels' FP32 inverse normal with the device's native log2
    and sqrt, over every 32-bit word, against the FP64 AS241 of the same
    midpoint probability: relative error of z below 7e-7."""
    import ctypes as C

    err, word = C.c_double(), C.c_uint32()
    assert engine.load().pe_normal_mixed_max_error(0, 1 << 32, C.byref(err), C.byref(word)) == abi.PE_OK
    print("max relative error %.3e at word %d" % (err.value, word.value), flush=True)
    assert err.value < 7e-7, (err.value, word.value)
