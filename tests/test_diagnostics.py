"""MonteCarloDiagnostics carried across the C ABI (host side, no GPU).

* The library's BLAKE3-256 equals the wire module's independent BLAKE3 (both
  written from the specification) on lengths around the block, chunk and
  tree boundaries, and on the empty input's published digest.
* Every replay fixture's recorded ``monte_carlo`` block -- master seed,
  estimator, scramble count, the BLAKE3 direction and scramble checksums,
  policy versions, worker threads, reduction block, AAD tile capacity,
  checkpoint interval, antithetic flag and curve regions -- equals what
  pe_plan_diagnostics reports for the same request and execution policy
  (fixtures/replay/*-windows-x86_64.json, execution.monte_carlo), including
  the early-exercise cases' valuation engine.
"""
import ctypes as C

import pytest

from _support import golden
from pricing_amd import abi, engine, request, wire

G = golden()


def lib_blake3(data):
    out = (C.c_uint8 * 32)()
    buf = C.create_string_buffer(bytes(data), len(data)) if data else None
    assert engine.load().pe_blake3_256(buf, len(data), out) == abi.PE_OK
    return bytes(out).hex()


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 2049, 3072, 4096, 5 * 1024 + 7,
                               8 * 1024, 8 * 1024 + 1, 31 * 1024 + 999, 100_003])
def test_blake3_matches_the_wire_implementation(n):
    data = bytes((i * 251 + 7) % 256 for i in range(n))
    assert lib_blake3(data) == wire.blake3_256(data)


def test_blake3_empty_input_vector():
    assert lib_blake3(b"") == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"


def _diagnostics(doc, workers, block):
    req = request.Request.from_v3(doc)
    plan = engine.PricingPlan.compile(req, request.exec_policy(workers, block))
    d = abi.pe_mc_diagnostics()
    assert engine.load().pe_plan_diagnostics(plan._handle, C.byref(d)) == abi.PE_OK
    return d


def _check(d, mc):
    assert d.master_seed == int(mc["master_seed"])
    assert d.estimator == (abi.PE_ESTIMATOR_RQMC if mc["estimator"] == "randomized_quasi_monte_carlo"
                           else abi.PE_ESTIMATOR_PSEUDO_MONTE_CARLO)
    assert d.scramble_count == (mc["scramble_count"] or 0)
    if mc["direction_checksum"] is None:
        assert not d.has_checksums
    else:
        assert d.has_checksums
        assert bytes(d.direction_checksum).hex() == mc["direction_checksum"]
        assert bytes(d.scramble_checksum).hex() == mc["scramble_checksum"]
    assert d.policy_version == mc["policy_version"]
    assert d.worker_threads == mc["worker_threads"]
    assert d.reduction_block_size == mc["reduction_block_size"]
    assert d.aad_tile_policy_version == mc["aad_tile_policy_version"]
    assert d.aad_tile_capacity == mc["aad_tile_capacity"]
    assert d.checkpoint_policy_version == mc["checkpoint_policy_version"]
    assert d.checkpoint_interval == mc["checkpoint_interval"]
    assert bool(d.antithetic) == mc["antithetic"]
    assert abi.CURVE_REGIONS[d.discount_region] == mc["discount_region"]
    assert abi.CURVE_REGIONS[d.dividend_region] == mc["dividend_region"]


@pytest.mark.parametrize("case", G["replay"], ids=lambda c: c["source"].split("/")[-1].split("-")[0] + ":" + c["name"])
def test_replay_monte_carlo_block(case):
    d = _diagnostics(case["request"], case["worker_threads"], case["reduction_block_size"])
    _check(d, case["monte_carlo"])


@pytest.mark.parametrize("case", G["early_exercise_replay"], ids=lambda c: c["name"])
def test_early_exercise_monte_carlo_block(case):
    d = _diagnostics(case["request"], case["worker_threads"], case["reduction_block_size"])
    _check(d, case["monte_carlo"])


def test_a_tile_above_the_block_is_rejected():
    doc = next(c for c in G["replay"] if c["name"] == "pseudo_mc_full_risk")["request"]
    req = request.Request.from_v3(doc)
    req.c.risk.aad_tile_capacity = 512
    with pytest.raises(engine.EngineError) as err:
        engine.PricingPlan.compile(req, request.exec_policy(2, 256))
    assert err.value.status == abi.PE_ERR_INVALID_ARGUMENT
