"""GPU parity of least-squares Monte Carlo (American / Bermudan exercise):
training paths, backward induction and out-of-sample valuation on the device
against the CPU oracle and the reference's frozen American put replays
(fixtures/replay/early_exercise-windows-x86_64.json).

Bars: the FP64 device mode reproduces the oracle's price to 1e-12 relative,
its exercise counts exactly and its in-sample value to 1e-12; decision models
agree to the conditioning of the normal equations (the device solves the
Gram system with CPQR's pivot and rank rules, the oracle runs the reference's
Householder QR).  The mixed mode stays within 0.1 standard errors of FP64.
"""
import math

import pytest

from _support import golden, oracle_lsm
from pricing_amd import abi, engine, request, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


REPLAY = golden()["early_exercise_replay"]


def _device(req, backend="fp64", stopping_capacity=0):
    plan = engine.PricingPlan.compile(req)
    res, rep = engine.evaluate_lsm([plan], backend, stopping_capacity=stopping_capacity)
    return res[0], rep[0]


def _laguerre_prediction(coef, spot, strike):
    import ctypes as C
    from _support import oracle
    phi = (C.c_double * 4)()
    oracle().oracle_weighted_laguerre(spot / strike, 3, phi)
    return sum(a * b for a, b in zip(phi, coef))


def _compare_oracle(req, block=256):
    laguerre = req.c.lsm.method == abi.PE_LSM_POLY_LAGUERRE
    strike = req.c.product.strike
    dev, rep = _device(req)
    ref, diag, ins, counts = oracle_lsm(req, block)
    n = rep.exercise_dates
    assert rep.exercise_counts == counts[:n]
    assert abs(dev.estimates[0].value - ref.estimates[0].value) <= 1e-12 * abs(ref.estimates[0].value)
    assert abs(dev.estimates[0].standard_error - ref.estimates[0].standard_error) <= \
        1e-10 * ref.estimates[0].standard_error
    assert abs(rep.in_sample_value - ins) <= 1e-12 * abs(ins)
    for d, row in enumerate(diag):
        m = rep.decisions[d]
        assert m.continue_all == int(row[19]) and m.itm_rows == int(row[0]), d
        if m.continue_all:
            continue
        assert m.rank == int(row[1]), d
        assert list(m.pivot_order)[:m.rank] == [int(v) for v in row[14:14 + m.rank]], d
        if laguerre:
            # the weighted Laguerre Gram is ill-conditioned on the in-the-money
            # range (condition ~1e9: the summation order alone moves the fit
            # by ~1e-8): compare the fitted continuation inside the data range
            for spot in (70.0, 80.0, 90.0, 95.0, 99.0):
                a = _laguerre_prediction(m.coefficients, spot, strike)
                b = _laguerre_prediction(row[6:10], spot, strike)
                assert abs(a - b) <= 1e-6 * max(1.0, abs(b)), (d, spot, a, b)
            for k in range(4):
                assert abs(m.diagonal_abs[k] - row[10 + k]) <= 1e-7 * row[10], (d, k)
            continue
        for k in range(4):
            assert abs(m.coefficients[k] - row[6 + k]) <= 1e-8 * max(1.0, abs(row[6 + k])), (d, k)
            assert abs(m.diagonal_abs[k] - row[10 + k]) <= 1e-9 * max(row[10 + k], 1e-300), (d, k)
    return dev, rep


@pytest.mark.parametrize("case", REPLAY, ids=[c["name"] for c in REPLAY])
def test_replay_on_device(case):
    req = request.Request.from_v3(case["request"])
    dev, rep = _device(req, stopping_capacity=len(case["early_exercise"]["stopping_indices"]))
    exp = case["value"]
    # fixed-policy Greeks: pathwise at each path's stopping date; Gamma is a
    # difference of pathwise Deltas over 2h (rounding scale |Delta| / h)
    h = req.c.market.spot * req.c.risk.gamma_bump if req.c.risk.gamma else 1.0
    for key, idx, scale in (("delta", 1, 0.0), ("vega", 2, 0.0), ("gamma", 3, abs(dev.estimates[1].value) / h)):
        if key not in case["risks"]:
            continue
        e = case["risks"][key]
        assert abs(dev.estimates[idx].value - e["value"]) <= 1e-12 * (abs(e["value"]) + scale), (key, dev.estimates[idx].value, e["value"])
        assert abs(dev.estimates[idx].standard_error - e["standard_error"]) <= 1e-10 * e["standard_error"], key
    assert abs(dev.estimates[0].value - exp["value"]) <= 1e-12 * abs(exp["value"])
    assert abs(dev.estimates[0].standard_error - exp["standard_error"]) <= 1e-10 * exp["standard_error"]
    ee = case["early_exercise"]
    assert rep.exercise_counts[:len(ee["exercise_counts"])] == ee["exercise_counts"]
    # EarlyExerciseDiagnostics: probabilities, stopping index of every
    # valuation trajectory, collisions, engine sizes, domains and checksums
    assert rep.exercise_probabilities == ee["exercise_probabilities"]
    assert rep.dividend_collisions == ee["dividend_collisions"]
    assert rep.stopping_indices == ee["stopping_indices"]
    assert rep.training_sampling_units == ee["training_sampling_units"]
    assert rep.training_trajectories == ee["training_trajectories"]
    assert rep.valuation_sampling_units == ee["valuation_sampling_units"]
    assert rep.valuation_trajectories == ee["valuation_trajectories"]
    assert abi.RANDOM_DOMAINS[rep.training_random_domain] == ee["training_random_domain"]["type"]
    assert abi.RANDOM_DOMAINS[rep.valuation_random_domain] == ee["valuation_random_domain"]["type"]
    for k in ("training_direction_checksum", "training_scramble_checksum", "valuation_direction_checksum",
              "valuation_scramble_checksum"):
        assert getattr(rep, k) == ee.get(k), k
    assert rep.policy_max_degree == ee["policy_basis"]["max_degree"]
    assert rep.policy_feature_count == ee["policy_basis"]["feature_count"]
    assert rep.itm_abs_tolerance == ee["itm_abs_tolerance"] and rep.max_matrix_elements == ee["max_matrix_elements"]
    assert rep.abs_rank_tolerance == ee["cpqr"]["abs_rank_tolerance"]
    assert rep.rel_rank_tolerance == ee["cpqr"]["rel_rank_tolerance"]
    mc = dev.diagnostics
    assert mc.master_seed == int(case["monte_carlo"]["master_seed"]) and mc.policy_version == 1
    assert abs(rep.in_sample_value - ee["in_sample_value"]) <= 1e-12 * abs(ee["in_sample_value"])
    for d, model in enumerate(ee["decision_models"]):
        m = model["model"]
        got = rep.decisions[d]
        assert got.rank == m["rank"]
        assert list(got.pivot_order)[:len(m["pivot_order"])] == m["pivot_order"]
        for k in range(4):
            assert abs(got.coefficients[k] - m["coefficients"][k]) <= 1e-9 * max(1.0, abs(m["coefficients"][k]))
            assert abs(got.diagonal_abs[k] - m["diagonal_abs"][k]) <= 1e-10 * m["diagonal_abs"][k]
        sc = m["feature_scalings"][0]
        assert abs(got.feature_mean - sc["mean"]) <= 1e-13 * abs(sc["mean"])
        assert abs(got.feature_population_variance - sc["population_variance"]) <= 1e-11 * sc["population_variance"]
        assert abs(got.residual_sum_squares - m["residual_sum_squares"]) <= 1e-8 * m["residual_sum_squares"]


def _put(method, units, training_units, vol=0.25, rqmc=False, divs=(), antithetic=True, dates=None, strike=100.0):
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-0.05)], [0.0, 1.0], [1.0, 1.0],
                                       list(divs))
    req.set_black_scholes(vol)
    dates = dates or [73, 146, 219, 292, 365]
    req.set_product(abi.PE_PRODUCT_AMERICAN, abi.PE_PUT, dates[-1], strike, dates=dates)
    if rqmc:
        req.set_rqmc(0xFEDCBA9876543210, units, 4, antithetic=antithetic)
        req.set_lsm((abi.PE_ENGINE_RQMC, 0x1234, training_units, 4), antithetic=antithetic, method=method)
    else:
        req.set_pseudo(8738, units, antithetic=antithetic)
        req.set_lsm((abi.PE_ENGINE_PSEUDO, 4369, training_units, 0), antithetic=antithetic, method=method)
    req.set_risk()
    return req


METHODS = {"cpqr": abi.PE_LSM_POLY_CPQR, "laguerre": abi.PE_LSM_POLY_LAGUERRE}
MONTHLY = [30, 61, 91, 122, 152, 182, 213, 243, 274, 304, 335, 365]


@pytest.mark.parametrize("method", sorted(METHODS))
@pytest.mark.parametrize("shape", ["pseudo", "rqmc", "dividends", "monthly"])
def test_fp64_matches_oracle(method, shape):
    kw = {"pseudo": {}, "rqmc": {"rqmc": True}, "dividends": {"divs": [(0.5, 1.0, 0.01)]},
          "monthly": {"dates": MONTHLY, "antithetic": False}}[shape]
    req = _put(METHODS[method], 1024, 2048, **kw)
    print(method, shape, flush=True)
    _compare_oracle(req)


def test_zero_volatility_put_exercises_at_the_first_date():
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-0.05)], [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(0.0)
    req.set_product(abi.PE_PRODUCT_AMERICAN, abi.PE_PUT, 365, 120.0, dates=[91, 182, 274, 365])
    req.set_pseudo(1, 64)
    req.set_risk()
    for method in METHODS.values():
        req.set_lsm((abi.PE_ENGINE_PSEUDO, 2, 64, 0), method=method)
        for backend in ("fp64", "mixed"):
            dev, rep = _device(req, backend)
            s1 = 100.0 * math.exp(0.05 * 91 / 365)
            assert dev.estimates[0].value == pytest.approx((120.0 - s1) * math.exp(-0.05 * 91 / 365), rel=1e-6)
            assert rep.exercise_counts[0] == 64


@pytest.mark.parametrize("method", sorted(METHODS))
def test_mixed_within_tenth_se(method):
    req = _put(METHODS[method], 1 << 15, 1 << 15, dates=MONTHLY, antithetic=False)
    plan = engine.PricingPlan.compile(req)
    (f64,), _ = engine.evaluate_lsm([plan], "fp64")
    (mixed,), _ = engine.evaluate_lsm([plan], "mixed")
    se = f64.estimates[0].standard_error
    assert abs(mixed.estimates[0].value - f64.estimates[0].value) < 0.1 * se, (
        mixed.estimates[0].value, f64.estimates[0].value, se)


def test_batch_equals_single_plans_and_oracle():
    reqs = workloads.american_batch(24, seed=5, train_log2=10, value_log2=10, dates=12, dividends=True)
    plans = [engine.PricingPlan.compile(r) for r in reqs]
    batch, reps = engine.evaluate_lsm(plans, "fp64")
    via_batch_api = engine.evaluate_batch(plans, "fp64")
    for i in (0, 5, 17):
        single, rep = engine.evaluate_lsm([plans[i]], "fp64")
        assert single[0].estimates[0].value == batch[i].estimates[0].value
        assert rep[0].exercise_counts == reps[i].exercise_counts
    for i, r in enumerate(reqs):
        assert via_batch_api[i].estimates[0].value == batch[i].estimates[0].value
        if i % 6 == 0:
            ref, _, ins, counts = oracle_lsm(r)
            assert reps[i].exercise_counts == counts[:12]
            assert abs(batch[i].estimates[0].value - ref.estimates[0].value) <= 1e-12 * ref.estimates[0].value


def test_laguerre_batch_mixed_within_tenth_se():
    """The full 256-trade x 50-date American workload at 2^18 training and
    2^18 valuation paths per trade: every mixed price within 0.1 SE of the
    FP64 mode.  The mixed mode trains on FP64 paths (decisions identical to
    the FP64 mode) and values in FP32."""
    reqs = workloads.american_batch(256, train_log2=18, value_log2=18, dates=50)
    plans = [engine.PricingPlan.compile(r) for r in reqs]
    f64, frep = engine.evaluate_lsm(plans, "fp64")
    mixed, mrep = engine.evaluate_lsm(plans, "mixed")
    assert [d.coefficients[:] for d in frep[7].decisions] == [d.coefficients[:] for d in mrep[7].decisions]
    worst = max(abs(a.estimates[0].value - b.estimates[0].value) / a.estimates[0].standard_error
                for a, b in zip(f64, mixed))
    print("max mixed - fp64 in SE: %.4f" % worst, flush=True)
    for r, a, b in zip(reqs, f64, mixed):
        se = a.estimates[0].standard_error
        assert abs(a.estimates[0].value - b.estimates[0].value) < 0.1 * se, (r.trade_id, a.estimates[0].value,
                                                                             b.estimates[0].value, se)


# ---- learned policy: one continuation network per trade, judged by Laguerre --

def _mlp_batch(method, train_log2, n=64, value_log2=13, dates=12, seed=21, backend="mixed"):
    reqs = workloads.american_batch(n, seed=seed, train_log2=train_log2, value_log2=value_log2, dates=dates,
                                    method=method)
    plans = [engine.PricingPlan.compile(r) for r in reqs]
    return engine.evaluate_lsm(plans, backend)


def test_mlp_policy_is_deterministic_with_valid_regions():
    a, rep_a = _mlp_batch(abi.PE_LSM_MLP, 9)
    b, rep_b = _mlp_batch(abi.PE_LSM_MLP, 9)
    for x, y in zip(a, b):
        assert x.estimates[0].value == y.estimates[0].value
    for r, s in zip(rep_a, rep_b):
        assert r.exercise_counts == s.exercise_counts
        assert r.network_price.value == s.network_price.value and r.judge_price.value == s.judge_price.value
        for c in r.decisions:
            assert c.region_cut_count <= 63
            cuts = list(c.region_cuts)[:c.region_cut_count]
            assert all(0.0 < x < 1.0 for x in cuts) and cuts == sorted(cuts)


def test_judge_is_the_laguerre_policy_and_the_fallback_keeps_the_higher_value():
    """In FP64 mode the judge is bit for bit the weighted-Laguerre plan on the
    same training and valuation paths; the result is the higher of the
    network's and the judge's out-of-sample values, and the exercise counts
    are the kept policy's."""
    nn, nrep = _mlp_batch(abi.PE_LSM_MLP, 10, n=16, backend="fp64")
    lag, lrep = _mlp_batch(abi.PE_LSM_POLY_LAGUERRE, 10, n=16, backend="fp64")
    for r, x, lr in zip(nrep, lag, lrep):
        assert r.judge_price.value == x.estimates[0].value
        assert r.judge_price.standard_error == x.estimates[0].standard_error
        assert r.judge_in_sample_value == lr.in_sample_value
    for res, r, lr in zip(nn, nrep, lrep):
        keep = r.judge_price if r.fallback_used else r.network_price
        assert res.estimates[0].value == keep.value == max(r.network_price.value, r.judge_price.value)
        assert r.fallback_used == (r.judge_price.value > r.network_price.value)
        if r.fallback_used:
            assert r.exercise_counts == lr.exercise_counts
        assert r.network_training_ms > 0.0 and r.judge_training_ms > 0.0


def test_mlp_policy_within_one_standard_error_of_the_judge():
    """64 trades, 2^14 training and 2^13 valuation paths each, 12 dates: the
    per-trade networks price within 1 SE of the per-trade weighted-Laguerre
    regressions on the same paths, trade by trade (measured at 2^14
    valuation paths: mean +0.03 SE, worst -0.53 / +0.67 SE)."""
    import numpy as np
    _, rep = _mlp_batch(abi.PE_LSM_MLP, 14)
    z = np.array([(r.network_price.value - r.judge_price.value) / r.judge_price.standard_error for r in rep])
    print("network - judge in SE: mean %.3f min %.3f max %.3f" % (z.mean(), z.min(), z.max()), flush=True)
    assert abs(z.mean()) < 0.5 and np.max(np.abs(z)) < 1.0, z


def test_mlp_policy_gains_over_laguerre_on_few_training_paths():
    """What the network gains: on 2^8 training paths per trade (50 dates) the
    networks -- the European continuation as a control variate plus a trained
    residual -- value the 64 trades above their own per-trade Laguerre
    regressions on the same paths (measured: +2.3 SE on average)."""
    import numpy as np
    _, rep = _mlp_batch(abi.PE_LSM_MLP, 8, dates=50, value_log2=14)
    z = np.array([(r.network_price.value - r.judge_price.value) / r.judge_price.standard_error for r in rep])
    print("network - judge in SE at 2^8 training paths: mean %.3f" % z.mean(), flush=True)
    assert z.mean() > 1.0, z.mean()


def test_mlp_policy_captures_the_early_exercise_premium():
    """A single Bermudan put (the network's volatility input is then
    constant): the out-of-sample value of the trained network's policy is a
    lower bound within 3 SE of the binomial value and captures most of the
    premium over the European put."""
    from test_oracle_lsm import crr_bermudan_put
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-0.05)], [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(0.2)
    req.set_product(abi.PE_PRODUCT_AMERICAN, abi.PE_PUT, 365, 100.0, dates=[73, 146, 219, 292, 365])
    req.set_pseudo(8738, 1 << 15, antithetic=True)
    req.set_risk()
    req.set_lsm((abi.PE_ENGINE_PSEUDO, 4369, 1 << 14, 0), method=abi.PE_LSM_MLP)
    (res,), (rep,) = engine.evaluate_lsm([engine.PricingPlan.compile(req)], "mixed")
    v, se = res.estimates[0].value, res.estimates[0].standard_error
    tree = crr_bermudan_put(100.0, 100.0, 0.05, 0.2, 1.0, [0.2, 0.4, 0.6, 0.8, 1.0])
    european = 5.573526022256971
    assert v < tree + 3.0 * se
    assert v - european > 0.6 * (tree - european), (v, tree)


def _bump_scale(req, ref, c):
    """Rounding scale of component c: the bump-validation components divide
    price differences by the spot bump h (h^2) or the vol bump."""
    price = abs(ref.estimates[0].value) + 1e-300
    spot = req.c.market.spot
    h = spot * req.c.risk.gamma_bump if req.c.risk.gamma else spot * 1e-4
    vb = min(1e-4, 0.5 * req.c.model.volatility)
    extra = {3: abs(ref.estimates[1].value) / h, 4: price / h, 5: price / h, 6: price / vb, 7: price / vb,
             8: price / h ** 2, 9: price / h ** 2}
    return abs(ref.estimates[c].value) + extra.get(c, 0.0)


FIXED_POLICY = [
    ("pseudo-cpqr", dict(method=abi.PE_LSM_POLY_CPQR, units=2048, training_units=2048)),
    ("rqmc-laguerre", dict(method=abi.PE_LSM_POLY_LAGUERRE, units=512, training_units=512, rqmc=True)),
    ("dividends-cpqr", dict(method=abi.PE_LSM_POLY_CPQR, units=2048, training_units=2048,
                            divs=[(0.4, 1.0, 0.01)])),
]


@pytest.mark.parametrize("name,kw", FIXED_POLICY, ids=[f[0] for f in FIXED_POLICY])
def test_fixed_policy_greeks_fp64_match_oracle(name, kw):
    req = _put(**kw)
    req.set_risk(delta=True, vega=True, gamma_bump=0.01)
    print("case", name, flush=True)
    dev, _ = _device(req)
    ref = oracle_lsm(req)[0]
    for c in range(10):
        scale = _bump_scale(req, ref, c)
        assert abs(dev.estimates[c].value - ref.estimates[c].value) <= 1e-12 * scale + 1e-15, (
            c, dev.estimates[c].value, ref.estimates[c].value)
        se_ref = ref.estimates[c].standard_error
        se_scale = se_ref * scale / max(abs(ref.estimates[c].value), 1e-300)
        assert abs(dev.estimates[c].standard_error - se_ref) <= 1e-10 * max(se_ref, se_scale) + 1e-15, c


def test_fixed_policy_greeks_mixed_within_tenth_se():
    req = _put(abi.PE_LSM_POLY_LAGUERRE, 1 << 14, 1 << 14)
    req.set_risk(delta=True, vega=True, gamma_bump=0.01)
    plan = engine.PricingPlan.compile(req)
    f64 = engine.evaluate_lsm([plan], "fp64")[0][0]
    mixed = engine.evaluate_lsm([plan], "mixed")[0][0]
    for c in (0, 1, 2, 3):
        se = f64.estimates[c].standard_error
        assert abs(mixed.estimates[c].value - f64.estimates[c].value) <= 0.1 * se, (
            c, mixed.estimates[c].value, f64.estimates[c].value, se)


def test_fixed_policy_delta_matches_the_tree_delta():
    # The fixed-policy pathwise Delta of a near-optimal policy approximates the
    # Bermudan Delta (the policy's first-order effect vanishes at the optimum);
    # CRR tree Delta by central difference of tree prices.
    from test_oracle_lsm import crr_bermudan_put

    dates = [73, 146, 219, 292, 365]
    req = _put(abi.PE_LSM_POLY_LAGUERRE, 1 << 16, 1 << 16, vol=0.25, dates=dates)
    req.set_risk(delta=True)
    dev, _ = _device(req)
    times = [d / 365.0 for d in dates]
    up = crr_bermudan_put(100.5, 100.0, 0.05, 0.25, 1.0, times, steps=2000)
    dn = crr_bermudan_put(99.5, 100.0, 0.05, 0.25, 1.0, times, steps=2000)
    tree_delta = (up - dn) / 1.0
    assert abs(dev.estimates[1].value - tree_delta) < 4.0 * dev.estimates[1].standard_error + 0.005, (
        dev.estimates[1].value, tree_delta)
