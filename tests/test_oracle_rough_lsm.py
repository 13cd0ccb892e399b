"""Pins the oracle's early exercise on lifted-Heston paths: the reference's
LSM regression -- standardized monomials of spot and variance of total
degree <= 2, CPQR pivot and rank rules -- on the rough scheme's paths,
exercising in-the-money paths only (Longstaff and Schwartz; the reference's
all-path rule stops out-of-the-money paths wherever the quadratic fit
extrapolates below zero and loses 10 SE here).  No reference counterpart:
the reference rejects early exercise in its hybrid adapters.

* One exercise date: the policy has no decision and the price equals the
  oracle's lifted-Heston European put on the same valuation paths.
* Zero vol-of-vol and mean reversion: the variance stays v0, the variance
  feature is inactive (its columns pre-excluded), and the Bermudan put is
  the Black-Scholes one: within 4 SE (+ LSM bias allowance) of a CRR tree.
* At the smile model's parameters the Bermudan put is worth at least the
  European put on the same paths, exercises before expiry on every date's
  share of paths, and its regressions keep both features active.
"""
import math

from _support import oracle_rough, oracle_rough_lsm
from pricing_amd import abi, workloads
from test_oracle_lsm import crr_bermudan_put
from test_oracle_rough import from_rough

STEP = 2.0 / 256


def _smile_model():
    return workloads.lifted_heston_model(from_rough(0.1, 20, 2.5))


def _european_twin(req):
    """The same valuation engine and market with a European put at expiry."""
    from pricing_amd import request
    e = request.Request()
    e.c = abi.pe_request.from_buffer_copy(req.c)
    e._keep = req._keep
    e.set_product(abi.PE_PRODUCT_EUROPEAN, abi.PE_PUT, req.c.product.expiry_day, req.c.product.strike)
    return e


def test_single_exercise_date_is_the_european():
    model = _smile_model()
    req = workloads.rough_bermudan(units=1 << 12, train_units=1 << 10, dates=1, antithetic=True)
    am, rows, _, counts, _ = oracle_rough_lsm(req, model, STEP, workers=8)
    assert rows == [] and counts[0] == 2 * (1 << 12)
    eu, _ = oracle_rough([_european_twin(req)], model, STEP, workers=8)
    a, b = am.estimates[0], eu[0].estimates[0]
    assert abs(a.value - b.value) <= 1e-12 * abs(b.value), (a.value, b.value)
    assert abs(a.standard_error - b.standard_error) <= 1e-10 * b.standard_error


def test_constant_variance_limit_is_the_black_scholes_bermudan():
    v0 = 0.04
    w, x = from_rough(0.1, 20, 2.5)
    model = abi.lifted_heston(v0, 0.0, v0, 0.0, -0.7, w, x)
    req = workloads.rough_bermudan(units=1 << 15, train_units=1 << 15, dates=5, expiry_days=365)
    res, rows, _, counts, _ = oracle_rough_lsm(req, model, STEP, workers=8)
    for r in rows:
        assert r[28] == 1.0 and r[27] == 0.0, r   # variance inactive, spot active
        assert all(r[8 + k] == 0.0 for k in (2, 4, 5)), r   # columns with a variance power
    days = [max(1, int(round(365 * (k + 1) / 5))) for k in range(5)]
    tree = crr_bermudan_put(100.0, 100.0, workloads.BERMUDAN_RATE, math.sqrt(v0), days[-1] / 365.0,
                            [d / 365.0 for d in days])
    est = res.estimates[0]
    # LSM is a low-biased policy value: allow 4 SE plus a small bias
    assert tree - 4.0 * est.standard_error - 0.02 <= est.value <= tree + 4.0 * est.standard_error, (est.value, tree)


def test_smile_model_bermudan_exceeds_the_european_and_uses_both_features():
    model = _smile_model()
    req = workloads.rough_bermudan(units=1 << 14, train_units=1 << 14)
    res, rows, ins, counts, _ = oracle_rough_lsm(req, model, STEP, workers=8)
    eu, _ = oracle_rough([_european_twin(req)], model, STEP, workers=8)
    a, b = res.estimates[0], eu[0].estimates[0]
    assert a.value >= b.value - 3.0 * math.hypot(a.standard_error, b.standard_error), (a.value, b.value)
    assert sum(counts[:11]) > 0 and counts[0] > 0
    assert len(rows) == 11
    for r in rows:
        assert r[26] == 0.0 and r[27] == 0.0 and r[28] == 0.0 and r[1] >= 5, r
    assert abs(ins - a.value) < 0.05 * a.value
