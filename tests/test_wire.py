"""Schema-v3 wire against the reference's own documents: BLAKE3 and the
request fingerprint reproduce every replay fixture's recorded
request_fingerprint, and the pricing_result document built from the
oracle's result of a replay request reproduces the fixture's value and
risk blocks (raw and market-scaled units)."""
import json

import pytest

from _support import golden, oracle_evaluate
from pricing_amd import request, wire

G = golden()
CASES = G["replay"]


def test_blake3_empty_input():
    # BLAKE3 of the empty string, from the published specification
    assert wire.blake3_256(b"") == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"


def test_blake3_multi_chunk_tree_is_order_sensitive():
    data = bytes(i % 251 for i in range(5 * 1024))   # five distinct chunks, a three-level tree
    a = wire.blake3_256(data)
    b = wire.blake3_256(data[1024:2048] + data[:1024] + data[2048:])
    assert a != b and len(a) == 64


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_request_fingerprint_matches_the_fixture(case):
    assert wire.request_fingerprint(case["request"]) == case["request_fingerprint"]


def test_american_request_fingerprints_match():
    for case in golden()["early_exercise_replay"]:
        assert wire.request_fingerprint(case["request"]) == case["request_fingerprint"], case["name"]


def test_a_changed_field_changes_the_fingerprint():
    case = CASES[0]
    doc = json.loads(json.dumps(case["request"]))
    doc["market"]["spot"] = doc["market"]["spot"] + 1e-12
    assert wire.request_fingerprint(doc) != case["request_fingerprint"]


def _close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-15


@pytest.mark.parametrize("case", [c for c in CASES if c["name"] != "rqmc_delta_gamma_vega_vegakt"],
                         ids=[c["name"] for c in CASES if c["name"] != "rqmc_delta_gamma_vega_vegakt"])
def test_result_document_reproduces_the_fixture(case):
    req = request.Request.from_v3(case["request"])
    res = oracle_evaluate(req, workers=4, block=case["reduction_block_size"])
    doc = wire.result_document(case["request"], res)
    assert doc["replay"]["request_fingerprint"] == case["request_fingerprint"]
    exp = case["value"]
    got = doc["value"]
    assert got["estimator"] == exp["estimator"] and got["effective_sampling_units"] == exp["effective_sampling_units"]
    for k in ("value", "standard_error", "confidence_lower", "confidence_upper"):
        assert _close(got[k], exp[k], 1e-10 if k != "value" else 1e-12), (k, got[k], exp[k])
    assert set(doc["risks"]) == set(case["result_risks"])
    for name, block in case["result_risks"].items():
        mine = doc["risks"][name]
        for part in ("raw", "market_scaled"):
            assert mine[part + "_unit"] == block[part + "_unit"]
            for k in ("value", "standard_error", "confidence_lower", "confidence_upper"):
                assert _close(mine[part][k], block[part][k], 1e-10), (name, part, k, mine[part][k], block[part][k])


def test_result_document_has_the_schema_fields():
    case = CASES[0]
    res = oracle_evaluate(request.Request.from_v3(case["request"]), workers=4, block=case["reduction_block_size"])
    doc = wire.result_document(case["request"], res)
    golden_doc = golden()["v3_documents"]["pricing_result.golden.json"]
    assert set(golden_doc) <= set(doc) | {"monte_carlo"}
    assert set(doc) <= {"document_kind", "schema_version", "value", "risks", "diagnostics", "monte_carlo", "replay"}
    assert set(doc["replay"]) == set(golden_doc["replay"])


def test_v3_golden_american_documents_pair_up():
    # fixtures/v3: the American result document records the fingerprint of
    # the American request document (the European result holds a zero
    # placeholder fingerprint, so only the American pair pins it)
    v = golden()["v3_documents"]
    fp = wire.request_fingerprint(v["pricing_request_american.golden.json"])
    replay = v["pricing_result_american.golden.json"]["replay"]
    assert fp == replay["request_fingerprint"]
    assert fp == replay["migration"]["pre_migration_fingerprint"] == replay["migration"]["post_migration_fingerprint"]


def test_integer_written_floats_keep_the_fingerprint():
    """RequestV3 parses 100 and 100.0 into the same f64: the fingerprint of
    the canonical re-serialisation does not change (golden American request)."""
    import copy

    doc = G["v3_documents"]["pricing_request_american.golden.json"]
    want = wire.request_fingerprint(doc)
    alt = copy.deepcopy(doc)
    alt["market"]["spot"] = int(alt["market"]["spot"])
    alt["product"]["strike"] = int(alt["product"]["strike"])
    assert isinstance(alt["market"]["spot"], int)
    assert wire.request_fingerprint(alt) == want
    # an integer field stays an integer: changing its value changes the hash
    alt["engine"]["master_seed"] = int(alt["engine"]["master_seed"]) + 1
    assert wire.request_fingerprint(alt) != want


def test_graph_products_are_refused():
    import copy

    doc = copy.deepcopy(G["v3_documents"]["pricing_request.golden.json"])
    doc["product"] = {"type": "graph"}
    with pytest.raises(ValueError):
        wire.request_fingerprint(doc)
