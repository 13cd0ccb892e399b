"""Schema-v3 result documents built from device results.

* Every pricing replay fixture (European, local vol, path dependence): the
  pricing_result document built from the device's FP64 result reproduces the
  fixture's value and risk blocks (raw and market-scaled) and records the
  fixture's request fingerprint.
* The golden American documents (fixtures/v3/pricing_request_american and
  pricing_result_american): the document built from the device's FP64 result
  and early-exercise report reproduces the golden result's value, its
  monte_carlo.diagnostics block and its monte_carlo.early_exercise block --
  decision models, regression diagnostics, exercise counts, probabilities,
  every valuation trajectory's stopping index, in-sample value, engines --
  except the payoff-graph and exercise-policy fingerprints, which this engine
  does not compute.
"""
import pytest

from _support import golden
from pricing_amd import engine, request, wire

pytestmark = pytest.mark.gpu

G = golden()
CASES = [c for c in G["replay"] if c["name"] not in ("pseudo_mc_delta_gamma_vega_vegakt",
                                                      "rqmc_delta_gamma_vega_vegakt")]


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def _close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-15


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_device_result_document_reproduces_the_fixture(case):
    req = request.Request.from_v3(case["request"])
    plan = engine.PricingPlan.compile(req, request.exec_policy(case["worker_threads"], case["reduction_block_size"]))
    doc = wire.result_document(case["request"], plan.evaluate("fp64"))
    assert doc["replay"]["request_fingerprint"] == case["request_fingerprint"]
    exp, got = case["value"], doc["value"]
    assert got["estimator"] == exp["estimator"] and got["effective_sampling_units"] == exp["effective_sampling_units"]
    for k in ("value", "standard_error", "confidence_lower", "confidence_upper"):
        assert _close(got[k], exp[k], 1e-12 if k == "value" else 1e-10), (k, got[k], exp[k])
    assert set(doc["risks"]) == set(case["result_risks"])
    for name, block in case["result_risks"].items():
        for part in ("raw", "market_scaled"):
            assert doc["risks"][name][part + "_unit"] == block[part + "_unit"]
            for k in ("value", "standard_error"):
                assert _close(doc["risks"][name][part][k], block[part][k], 1e-10), (name, part, k)


def _compare(got, want, path, rel):
    if isinstance(want, dict):
        assert isinstance(got, dict) and set(got) == set(want), (path, sorted(set(got) ^ set(want)))
        for k in want:
            _compare(got[k], want[k], f"{path}.{k}", rel)
    elif isinstance(want, list):
        assert isinstance(got, list) and len(got) == len(want), (path, len(got), len(want))
        for i, (g, w) in enumerate(zip(got, want)):
            _compare(g, w, f"{path}[{i}]", rel)
    elif isinstance(want, bool) or isinstance(want, str) or want is None:
        assert got == want, (path, got, want)
    elif isinstance(want, int):
        assert got == want and not isinstance(got, float), (path, got, want)
    else:
        tol = rel.get(path.split(".")[-1].split("[")[0], 1e-12)
        assert _close(float(got), float(want), tol), (path, got, want)


def test_golden_american_result_document_on_device():
    req_doc = G["v3_documents"]["pricing_request_american.golden.json"]
    want = G["v3_documents"]["pricing_result_american.golden.json"]
    req = request.Request.from_v3(req_doc)
    mc = want["monte_carlo"]["diagnostics"]
    plan = engine.PricingPlan.compile(req, request.exec_policy(mc["worker_threads"], mc["reduction_block_size"]))
    ee = want["monte_carlo"]["early_exercise"]
    res, rep = engine.evaluate_lsm([plan], "fp64", stopping_capacity=len(ee["stopping_indices"]))
    doc = wire.result_document(req_doc, res[0], report=rep[0])
    assert doc["replay"]["request_fingerprint"] == want["replay"]["request_fingerprint"]
    for k in ("value", "standard_error", "confidence_lower", "confidence_upper"):
        assert _close(doc["value"][k], want["value"][k], 1e-12 if k == "value" else 1e-10), k
    assert doc["risks"] == want["risks"] == {}
    # the fitted regressions: coefficients of standardized monomials to 1e-9,
    # |R_pp| to 1e-10, the residual sum of squares to 1e-8 (an FP64 Gram +
    # pivoted Cholesky against the reference's Householder CPQR)
    rel = {"coefficients": 1e-9, "diagonal_abs": 1e-10, "residual_sum_squares": 1e-8, "rank_threshold": 1e-10,
           "population_variance": 1e-11, "scale": 1e-11, "mean": 1e-13}
    got_mc = doc["monte_carlo"]
    want_diag = {k: v for k, v in mc.items() if k != "payoff_fingerprint"}
    _compare(got_mc["diagnostics"], want_diag, "diagnostics", rel)
    want_ee = {k: v for k, v in ee.items() if k != "policy_fingerprint"}
    _compare(got_mc["early_exercise"], want_ee, "early_exercise", rel)
