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
