"""Pins the oracle's least-squares Monte Carlo exercise policy.

* Decimal reference cases of the early_exercise_v1 policy
  (fixtures/early-exercise/reference-cases-v0.1.json): exercise decisions on
  ties, polynomial basis enumeration, feature scaling and column-pivoted QR.
* The reference's frozen American put replays
  (fixtures/replay/early_exercise-windows-x86_64.json): price and SE, the
  fitted decision models (coefficients, QR diagonals, pivots, scalings,
  residual sums), exercise counts and the in-sample value.
* A deterministic limit: with zero volatility the optimal exercise of a put
  in the money is immediate when rates are positive.
"""
import ctypes as C
import math

import pytest

from _support import darr, golden, oracle
from pricing_amd import abi, request

G = golden()
CASES = G["early_exercise_cases"]


def test_decision_cases():
    for c in CASES["decision_cases"]:
        imm, cmp_ = float(c["immediate_value"]), float(c["comparison_value"])
        if c["kind"] == "exercise":
            assert (imm > cmp_) == c["expected"], c["id"]
        else:
            assert (imm > cmp_) == c["expected"], c["id"]


@pytest.mark.parametrize("case", CASES["basis_cases"], ids=lambda c: c["id"])
def test_basis_enumeration(case):
    out = (C.c_uint32 * 256)()
    n = oracle().oracle_basis_exponents(case["feature_count"], case["max_degree"], out, 256)
    fc = case["feature_count"]
    got = [list(out[i * fc:(i + 1) * fc]) for i in range(n)]
    assert got == case["expected_exponents"]


@pytest.mark.parametrize("case", CASES["scaling_cases"], ids=lambda c: c["id"])
def test_feature_scaling(case):
    vals = [float(v) for v in case["values"]]
    out = (C.c_double * 5)()
    oracle().oracle_feature_scaling(darr(vals), len(vals), out)
    e = case["expected"]
    for k, g in zip(("mean", "population_variance", "scale", "zero_scale_threshold"), out[:4]):
        assert abs(g - float(e[k])) <= 1e-15 * max(1.0, abs(float(e[k]))), k
    assert bool(out[4]) == e["inactive"]


@pytest.mark.parametrize("case", CASES["qr_cases"], ids=lambda c: c["id"])
def test_cpqr(case):
    m = [[float(x) for x in row] for row in case["matrix"]]
    rows, cols = len(m), len(m[0])
    flat = [x for row in m for x in row]
    t = [float(x) for x in case["target"]]
    out = (C.c_double * (2 + 2 * cols))()
    assert oracle().oracle_cpqr(darr(flat), rows, cols, darr(t), float(case["abs_rank_tol"]),
                                float(case["rel_rank_tol"]), out) == 0
    e = case["expected"]
    assert int(out[0]) == e["rank"]
    assert [int(out[2 + cols + c]) for c in range(cols)] == e["pivot_order"]
    for c in range(cols):
        assert abs(out[2 + c] - float(e["coefficients"][c])) <= 1e-13 * max(1.0, abs(float(e["coefficients"][c])))
    assert abs(out[1] - float(e["residual_sum_squares"])) <= 1e-12


REPLAY = G["early_exercise_replay"]


def lsm_evaluate(req, block=256):
    res = abi.pe_result()
    diag = (C.c_double * (20 * 64))()
    ins = C.c_double()
    counts = (C.c_uint64 * 64)()
    pol = request.exec_policy(4, block)
    n = oracle().oracle_lsm_evaluate(req.pointer(), C.byref(pol), C.byref(res), diag, 64, C.byref(ins), counts)
    assert n >= 0, oracle().oracle_last_error().decode()
    return res, [list(diag[20 * d:20 * d + 20]) for d in range(n)], ins.value, list(counts)


@pytest.mark.parametrize("case", REPLAY, ids=[c["name"] for c in REPLAY])
def test_american_replay(case):
    req = request.Request.from_v3(case["request"])
    res, diag, ins, counts = lsm_evaluate(req, case["reduction_block_size"])
    exp = case["value"]
    assert abs(res.estimates[0].value - exp["value"]) <= 1e-12 * abs(exp["value"])
    assert abs(res.estimates[0].standard_error - exp["standard_error"]) <= 1e-10 * exp["standard_error"]
    ee = case["early_exercise"]
    assert counts[:len(ee["exercise_counts"])] == ee["exercise_counts"]
    assert abs(ins - ee["in_sample_value"]) <= 1e-12 * abs(ee["in_sample_value"])
    for d, model in enumerate(ee["decision_models"]):
        m = model["model"]
        row = diag[d]
        assert int(row[1]) == m["rank"]
        assert [int(v) for v in row[14:18]] == m["pivot_order"]
        for k in range(4):
            assert abs(row[6 + k] - m["coefficients"][k]) <= 1e-11 * max(1.0, abs(m["coefficients"][k])), (d, k)
            assert abs(row[10 + k] - m["diagonal_abs"][k]) <= 1e-12 * m["diagonal_abs"][k]
        assert abs(row[4] - m["feature_scalings"][0]["mean"]) <= 1e-13 * abs(m["feature_scalings"][0]["mean"])
        assert abs(row[3] - m["residual_sum_squares"]) <= 1e-10 * m["residual_sum_squares"]
    # fixed-policy Greeks: pathwise at each path's stopping date; Gamma is a
    # difference of pathwise Deltas over 2h, so its rounding scale is |Delta| / h
    h = req.c.market.spot * req.c.risk.gamma_bump if req.c.risk.gamma else 1.0
    delta = res.estimates[1].value
    for key, idx, scale in (("delta", 1, 0.0), ("vega", 2, 0.0), ("gamma", 3, abs(delta) / h)):
        if key not in case["risks"]:
            continue
        e = case["risks"][key]
        assert abs(res.estimates[idx].value - e["value"]) <= 1e-12 * (abs(e["value"]) + scale), (key, res.estimates[idx].value, e["value"])
        assert abs(res.estimates[idx].standard_error - e["standard_error"]) <= 1e-10 * e["standard_error"], key


def test_zero_volatility_put_exercises_at_the_first_date():
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-0.05)], [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(0.0)
    req.set_product(abi.PE_PRODUCT_AMERICAN, abi.PE_PUT, 365, 120.0, dates=[91, 182, 274, 365])
    req.set_pseudo(1, 64)
    req.set_risk()
    req.set_lsm((abi.PE_ENGINE_PSEUDO, 2, 64, 0))
    res, diag, ins, counts = lsm_evaluate(req)
    s1 = 100.0 * math.exp(0.05 * 91 / 365)
    assert res.estimates[0].value == pytest.approx((120.0 - s1) * math.exp(-0.05 * 91 / 365), rel=1e-12)
    assert counts[0] == 64


# ---- normal equations + pivoted Cholesky (the device's regression) ----------

def _fit(fn, design, target, abs_tol=1e-14, rel_tol=1e-12, extra=0):
    rows, cols = len(design), len(design[0])
    flat = [x for row in design for x in row]
    out = (C.c_double * (2 + (2 + extra) * cols))()
    assert fn(darr(flat), rows, cols, darr(target), abs_tol, rel_tol, out) == 0, oracle().oracle_last_error()
    return int(out[0]), out[1], list(out[2:2 + cols]), [int(v) for v in out[2 + cols:2 + 2 * cols]], \
        list(out[2 + 2 * cols:2 + 3 * cols]) if extra else None


def _cpqr_diagonals(design, target):
    rows, cols = len(design), len(design[0])
    flat = [x for row in design for x in row]
    out = (C.c_double * cols)()
    assert oracle().oracle_cpqr_diagonals(darr(flat), rows, cols, darr(target), 1e-14, 1e-12, out) == 0
    return list(out)


def _standardized_monomials(xs):
    n = len(xs)
    mean = sum(xs) / n
    sd = math.sqrt(sum((x - mean) ** 2 for x in xs) / n)
    return [[1.0, s, s * s, s * s * s] for s in ((x - mean) / sd for x in xs)]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cholesky_fit_equals_reference_cpqr_full_rank(seed):
    import random
    rng = random.Random(seed)
    xs = [80.0 + 20.0 * rng.random() for _ in range(300)]
    design = _standardized_monomials(xs)
    target = [max(100.0 - x, 0.0) * 1.1 + rng.gauss(0.0, 2.0) for x in xs]
    rq, rssq, cq, pq, _ = _fit(oracle().oracle_cpqr, design, target)
    rc, rssc, cc, pc, dc = _fit(oracle().oracle_cholesky_fit, design, target, extra=1)
    assert rc == rq == 4 and pc == pq
    for a, b in zip(cc, cq):
        assert abs(a - b) <= 1e-10 * max(1.0, abs(b))
    for a, b in zip(dc, _cpqr_diagonals(design, target)):
        assert abs(a - b) <= 1e-11 * b
    assert abs(rssc - rssq) <= 1e-9 * rssq


def test_cholesky_fit_rank_deficient_keeps_the_cpqr_columns():
    # two distinct feature values: [1, s, s^2, s^3] has rank 2
    xs = [90.0] * 7 + [95.0] * 5
    design = _standardized_monomials(xs)
    target = [float(i % 3) for i in range(len(xs))]
    rq, _, cq, pq, _ = _fit(oracle().oracle_cpqr, design, target)
    rc, _, cc, pc, _ = _fit(oracle().oracle_cholesky_fit, design, target, extra=1)
    assert rc == rq == 2
    assert pc[:2] == pq[:2]
    for row in design:
        fq = sum(a * b for a, b in zip(row, cq))
        fc = sum(a * b for a, b in zip(row, cc))
        assert abs(fq - fc) <= 1e-9


def test_cholesky_fit_on_the_reference_qr_cases():
    """Well-conditioned Decimal QR cases of the reference: same rank, pivots
    and coefficients through the normal equations."""
    checked = 0
    for case in CASES["qr_cases"]:
        m = [[float(x) for x in row] for row in case["matrix"]]
        t = [float(x) for x in case["target"]]
        e = case["expected"]
        if len(m) < len(m[0]) or e["rank"] < len(m[0]):
            continue
        rc, _, cc, pc, _ = _fit(oracle().oracle_cholesky_fit, m, t, float(case["abs_rank_tol"]),
                                float(case["rel_rank_tol"]), extra=1)
        assert rc == e["rank"] and pc == e["pivot_order"], case["id"]
        for a, b in zip(cc, e["coefficients"]):
            assert abs(a - float(b)) <= 1e-10 * max(1.0, abs(float(b))), case["id"]
        checked += 1
    assert checked >= 1


def test_weighted_laguerre_closed_form():
    from scipy.special import eval_laguerre
    out = (C.c_double * 6)()
    for x in (0.0, 0.3, 0.9, 1.7, 4.2):
        oracle().oracle_weighted_laguerre(x, 5, out)
        for k in range(6):
            want = math.exp(-0.5 * x) * eval_laguerre(k, x)
            assert abs(out[k] - want) <= 1e-14 * max(1.0, abs(want)), (x, k)


def test_laguerre_fit_recovers_a_target_in_its_span():
    out = (C.c_double * 4)()
    coef = [3.0, -1.5, 0.25, 2.0]
    design, target = [], []
    for i in range(40):
        x = 0.5 + i / 80.0
        oracle().oracle_weighted_laguerre(x, 3, out)
        row = list(out)
        design.append(row)
        target.append(sum(a * b for a, b in zip(row, coef)))
    rank, rss, c, _, _ = _fit(oracle().oracle_cholesky_fit, design, target, extra=1)
    assert rank == 4
    for a, b in zip(c, coef):
        assert abs(a - b) <= 1e-6 * abs(b)   # Laguerre functions on [0.5, 1] are nearly dependent
    assert abs(rss) <= 1e-12


def _put_request(method, units=512, training_units=512, vol=0.2, seed=8738):
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-0.05)], [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(vol)
    req.set_product(abi.PE_PRODUCT_AMERICAN, abi.PE_PUT, 365, 100.0, dates=[73, 146, 219, 292, 365])
    req.set_pseudo(seed, units, antithetic=True)
    req.set_risk()
    req.set_lsm((abi.PE_ENGINE_PSEUDO, 4369, training_units, 0), method=method)
    return req


def crr_bermudan_put(spot, strike, rate, vol, expiry, exercise_times, steps=4000):
    """Cox-Ross-Rubinstein tree with exercise only at the given times."""
    import numpy as np
    dt = expiry / steps
    up = math.exp(vol * math.sqrt(dt))
    down = 1.0 / up
    p = (math.exp(rate * dt) - down) / (up - down)
    disc = math.exp(-rate * dt)
    ex = {int(round(t / dt)) for t in exercise_times}
    j = np.arange(steps + 1)
    v = np.maximum(strike - spot * up ** (steps - j) * down ** j, 0.0)
    for n in range(steps - 1, -1, -1):
        v = disc * (p * v[:-1] + (1.0 - p) * v[1:])
        if n in ex:
            k = np.arange(n + 1)
            v = np.maximum(v, strike - spot * up ** (n - k) * down ** k)
    return float(v[0])


def test_laguerre_policy_prices_the_bermudan_put():
    """Out-of-sample Laguerre policy (a lower-bound estimator) against the
    binomial Bermudan value: within 3 SE plus a small policy loss, and above
    the European put."""
    tree = crr_bermudan_put(100.0, 100.0, 0.05, 0.2, 1.0, [0.2, 0.4, 0.6, 0.8, 1.0])
    assert abs(tree - 5.9810) < 1e-3
    res, diag, ins, counts = lsm_evaluate(_put_request(abi.PE_LSM_POLY_LAGUERRE, 16384, 16384))
    v, se = res.estimates[0].value, res.estimates[0].standard_error
    assert tree - 3.0 * se - 0.03 < v < tree + 3.0 * se
    assert v > 5.573526022256971   # Black-Scholes European put, S=K=100, r=5%, sigma=20%, T=1
    assert all(int(r[1]) == 4 for r in diag)


def test_reference_policy_stops_out_of_the_money_paths():
    """The reference compares immediate and continuation values on every
    path: with few training paths the cubic's negative extrapolation beyond
    the strike stops out-of-the-money paths at zero, well below the
    in-the-money-only Laguerre policy on the same valuation paths."""
    a, _, _, _ = lsm_evaluate(_put_request(abi.PE_LSM_POLY_CPQR, 4096, 2048))
    b, _, _, _ = lsm_evaluate(_put_request(abi.PE_LSM_POLY_LAGUERRE, 4096, 2048))
    assert b.estimates[0].value - a.estimates[0].value > 5.0 * a.estimates[0].standard_error


def test_laguerre_zero_volatility_put_exercises_at_the_first_date():
    req = request.Request().set_market(100.0, [0.0, 1.0], [1.0, math.exp(-0.05)], [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(0.0)
    req.set_product(abi.PE_PRODUCT_AMERICAN, abi.PE_PUT, 365, 120.0, dates=[91, 182, 274, 365])
    req.set_pseudo(1, 64)
    req.set_risk()
    req.set_lsm((abi.PE_ENGINE_PSEUDO, 2, 64, 0), method=abi.PE_LSM_POLY_LAGUERRE)
    res, diag, ins, counts = lsm_evaluate(req)
    s1 = 100.0 * math.exp(0.05 * 91 / 365)
    assert res.estimates[0].value == pytest.approx((120.0 - s1) * math.exp(-0.05 * 91 / 365), rel=1e-12)
    assert counts[0] == 64
