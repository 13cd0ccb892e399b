"""Schema-v3 wire: request fingerprints and pricing_result documents.

* ``request_fingerprint(doc)`` -- the reference's ``fingerprint_request``
  (crates/pricing/src/wire.rs:4144-4212): the request value encoded with
  type tags and big-endian lengths (objects in key order), prefixed by
  ``b"pricing/request\\0"`` and the schema version, hashed with BLAKE3-256.
* ``result_document(request, result)`` -- a ``pricing_result`` document
  (schemas/v3/pricing_result.schema.json): the value estimate, the requested
  risks in raw and market-scaled units (build_risk_report,
  crates/pricing/src/engine/risk/report.rs:255-300, 475-489) and the replay
  block with the request fingerprint.

BLAKE3 is written here from its published specification (O'Connor, Aumasson,
Neves, Wilcox-O'Hearn, "BLAKE3", 2020): the ChaCha-derived compression
function, 1 KiB chunks and the binary chaining-value tree; unkeyed hash,
32-byte output.
"""
import struct

__all__ = ["blake3_256", "encode_value", "canonical_request", "request_fingerprint", "estimate_document",
           "monte_carlo_diagnostics_document", "early_exercise_document", "result_document"]

_IV = (0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19)
_PERMUTATION = (2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8)
_CHUNK_START, _CHUNK_END, _PARENT, _ROOT = 1, 2, 4, 8
_MASK = 0xFFFFFFFF
_CHUNK = 1024
_BLOCK = 64
NORMAL_95 = 1.959963984540054   # crates/pricing/src/engine/plan/simulation.rs:24


def _rotr(x, n):
    return ((x >> n) | (x << (32 - n))) & _MASK


def _g(s, a, b, c, d, x, y):
    s[a] = (s[a] + s[b] + x) & _MASK
    s[d] = _rotr(s[d] ^ s[a], 16)
    s[c] = (s[c] + s[d]) & _MASK
    s[b] = _rotr(s[b] ^ s[c], 12)
    s[a] = (s[a] + s[b] + y) & _MASK
    s[d] = _rotr(s[d] ^ s[a], 8)
    s[c] = (s[c] + s[d]) & _MASK
    s[b] = _rotr(s[b] ^ s[c], 7)


def _compress(cv, m, counter, block_len, flags):
    s = list(cv) + list(_IV[:4]) + [counter & _MASK, (counter >> 32) & _MASK, block_len, flags]
    m = list(m)
    for r in range(7):
        _g(s, 0, 4, 8, 12, m[0], m[1])
        _g(s, 1, 5, 9, 13, m[2], m[3])
        _g(s, 2, 6, 10, 14, m[4], m[5])
        _g(s, 3, 7, 11, 15, m[6], m[7])
        _g(s, 0, 5, 10, 15, m[8], m[9])
        _g(s, 1, 6, 11, 12, m[10], m[11])
        _g(s, 2, 7, 8, 13, m[12], m[13])
        _g(s, 3, 4, 9, 14, m[14], m[15])
        if r < 6:
            m = [m[i] for i in _PERMUTATION]
    for i in range(8):
        s[i] ^= s[i + 8]
        s[i + 8] ^= cv[i]
    return s


def _words(block):
    return struct.unpack("<16I", block + b"\0" * (_BLOCK - len(block)))


def _chunk_output(data, counter):
    """Chains a chunk's blocks; returns the compression inputs of its last block."""
    blocks = [data[i:i + _BLOCK] for i in range(0, len(data), _BLOCK)] or [b""]
    cv = _IV
    for i, block in enumerate(blocks):
        flags = (_CHUNK_START if i == 0 else 0) | (_CHUNK_END if i == len(blocks) - 1 else 0)
        if i == len(blocks) - 1:
            return cv, _words(block), counter, len(block), flags
        cv = _compress(cv, _words(block), counter, _BLOCK, flags)[:8]


def _chaining_value(output):
    return _compress(*output)[:8]


def _parent_output(left, right):
    return _IV, tuple(left) + tuple(right), 0, _BLOCK, _PARENT


def blake3_256(data):
    """BLAKE3 hash of `data` (bytes), 32 bytes as lowercase hex."""
    data = bytes(data)
    chunks = [data[i:i + _CHUNK] for i in range(0, len(data), _CHUNK)] or [b""]
    stack = []
    for index, chunk in enumerate(chunks[:-1]):
        cv = _chaining_value(_chunk_output(chunk, index))
        total = index + 1
        while total & 1 == 0:      # merge completed subtrees
            cv = _chaining_value(_parent_output(stack.pop(), cv))
            total >>= 1
        stack.append(cv)
    output = _chunk_output(chunks[-1], len(chunks) - 1)
    while stack:
        output = _parent_output(stack.pop(), _chaining_value(output))
    cv, m, counter, block_len, flags = output
    return struct.pack("<8I", *_compress(cv, m, counter, block_len, flags | _ROOT)[:8]).hex()


def encode_value(value, out=None):
    """encode_value (wire.rs:4165-4212): tag byte, u64 big-endian payload
    length, payload; integers as u64/i64, other numbers as f64 bits, objects
    as (key, value) pairs in key order."""
    out = bytearray() if out is None else out
    payload = bytearray()
    if value is None:
        tag = 0
    elif isinstance(value, bool):
        payload.append(1 if value else 0)
        tag = 1
    elif isinstance(value, int):
        if value >= 0:
            payload += struct.pack(">Q", value)
            tag = 2
        else:
            payload += struct.pack(">q", value)
            tag = 3
    elif isinstance(value, float):
        if value != value or value in (float("inf"), float("-inf")):
            raise ValueError("non-finite JSON number")
        payload += struct.pack(">d", value)
        tag = 4
    elif isinstance(value, str):
        payload += value.encode("utf-8")
        tag = 5
    elif isinstance(value, (list, tuple)):
        payload += struct.pack(">Q", len(value))
        for item in value:
            encode_value(item, payload)
        tag = 6
    elif isinstance(value, dict):
        payload += struct.pack(">Q", len(value))
        for key in sorted(value, key=lambda k: k.encode("utf-8")):
            encode_value(key, payload)
            encode_value(value[key], payload)
        tag = 7
    else:
        raise TypeError(f"not a JSON value: {type(value).__name__}")
    out.append(tag)
    out += struct.pack(">Q", len(payload))
    out += payload
    return out


# Fields the v3 request schema types as JSON numbers (schemas/v3/pricing_request.schema.json):
# RequestV3 holds them as f64, so its re-serialisation -- what fingerprint_request
# hashes (wire.rs:4144-4161) -- writes them as floats even when the input wrote an
# integer.  Every other number in the schema is an integer field.  No name is
# typed both ways, so the field name decides.
_FLOAT_FIELDS = frozenset({
    "abs_rank_tolerance", "amount", "barrier", "beta", "cap", "cdf_tolerance", "discount_factors", "ex_time",
    "fixed_cash", "fixing", "floor", "grid_width", "half_width", "historical_extremum", "implied_volatilities",
    "itm_abs_tolerance", "iv_bump", "log_forward_moneyness_nodes", "maturity_nodes", "notional",
    "payoff_smoothing_width_ladder", "payout", "projection_mean_tolerance", "projection_nodes",
    "projection_tail_tolerance", "rebate", "rel_rank_tolerance", "relative_density_threshold", "spot", "strike",
    "tail_tolerance", "time_nodes", "times", "valuation_date_fixing", "value", "values", "volatility", "weight",
})

# Product variants of the v3 request schema; a graph product has no v3
# request form (its fingerprint uses another domain) and is refused.
_V3_PRODUCTS = frozenset({"european_vanilla", "american_vanilla", "digital", "barrier", "arithmetic_asian",
                          "fixed_lookback"})


def _as_float(value):
    if isinstance(value, bool):
        return value
    if isinstance(value, int):
        return float(value)
    if isinstance(value, list):
        return [_as_float(v) for v in value]
    return value


def canonical_request(value, key=None):
    """The request document as RequestV3 re-serialises it: numbers of
    float-typed fields as floats (an integer 100 becomes 100.0)."""
    if isinstance(value, dict):
        return {k: canonical_request(v, k) for k, v in value.items()}
    if isinstance(value, list):
        return [canonical_request(v, key) for v in value]
    if key in _FLOAT_FIELDS:
        return _as_float(value)
    return value


def request_fingerprint(doc):
    """fingerprint_request for a schema-v3 pricing_request document, hashed
    in its canonical form (canonical_request)."""
    if doc.get("document_kind") != "pricing_request":
        raise ValueError("not a pricing_request document")
    product = doc.get("product", {}).get("type")
    if product not in _V3_PRODUCTS:
        raise ValueError(f"product {product!r} has no schema-v3 request fingerprint")
    data = bytearray(b"pricing/request\0") + struct.pack(">I", int(doc.get("schema_version", 3)))
    encode_value(canonical_request(doc), data)
    return "blake3-256:" + blake3_256(bytes(data))


def estimate_document(value, standard_error, effective_units, estimator, scale=1.0):
    """Estimate (estimate_from_statistics, report.rs:454-473) of a component
    scaled by `scale`: value x scale, SE x |scale|, 95% bracket."""
    v = value * scale
    se = standard_error * abs(scale)
    half = NORMAL_95 * se
    return {"value": v, "standard_error": se, "confidence_lower": v - half, "confidence_upper": v + half,
            "estimator": {"type": estimator}, "effective_sampling_units": int(effective_units)}


_REGIONS = ("pillar", "interpolated", "right_extrapolated")
_VALUATION_KINDS = ("exact_contractual", "smoothed_surrogate")
_DOMAINS = ("valuation", "lsm_train", "rqmc_scramble")


def monte_carlo_diagnostics_document(d):
    """MonteCarloDiagnostics (mc_result.rs:226-247) of a pe_result as the
    result document writes it; the payoff-graph fingerprint is not carried."""
    rqmc = d.estimator == 1
    out = {
        "aad_tile_capacity": int(d.aad_tile_capacity), "aad_tile_policy_version": int(d.aad_tile_policy_version),
        "antithetic": bool(d.antithetic), "checkpoint_interval": int(d.checkpoint_interval),
        "checkpoint_policy_version": int(d.checkpoint_policy_version),
        "discount_region": {"type": _REGIONS[d.discount_region]},
        "dividend_region": {"type": _REGIONS[d.dividend_region]},
        "estimator": {"type": "randomized_quasi_monte_carlo" if rqmc else "pseudo_monte_carlo"},
        "master_seed": int(d.master_seed), "policy_version": int(d.policy_version),
        "reduction_block_size": int(d.reduction_block_size),
        "valuation_kind": {"type": _VALUATION_KINDS[d.valuation_kind]}, "worker_threads": int(d.worker_threads),
    }
    if rqmc:
        out["scramble_count"] = int(d.scramble_count)
        out["direction_checksum"] = bytes(d.direction_checksum).hex()
        out["scramble_checksum"] = bytes(d.scramble_checksum).hex()
    return out


def _decision_document(m, columns, max_degree):
    """ExerciseDecisionModel of one date (wire.rs:2048-2090): a polynomial
    regression on the standardized spot, or continue-all without in-the-money
    training rows."""
    if m.continue_all:
        return {"type": "continue_all", "reason": {"type": "zero_itm_training_paths"}}, [
            {"type": "zero_itm_training_paths"}]
    pre_excluded = list(range(1, columns)) if m.feature_inactive else []
    active = [c for c in range(columns) if c not in pre_excluded]
    pivot = [int(v) for v in list(m.pivot_order)[:len(active)]]
    rank = int(m.rank)
    model = {
        "active_basis_columns": active,
        "basis": {"exponents": [[k] for k in range(columns)], "feature_count": 1, "max_degree": max_degree},
        "coefficients": [float(v) for v in list(m.coefficients)[:columns]],
        "diagonal_abs": [float(v) for v in list(m.diagonal_abs)[:len(active)]],
        "feature_scalings": [{"inactive": bool(m.feature_inactive), "mean": m.feature_mean,
                              "population_variance": m.feature_population_variance, "scale": m.feature_scale,
                              "zero_scale_threshold": m.zero_scale_threshold}],
        "pivot_order": pivot,
        "pre_excluded_basis_columns": pre_excluded,
        "rank": rank,
        "rank_excluded_basis_columns": pivot[rank:],
        "rank_threshold": m.rank_threshold,
        "residual_sum_squares": m.residual_sum_squares,
    }
    warnings = ([{"type": "inactive_feature", "feature": 0}] if m.feature_inactive else []) + [
        {"type": "rank_excluded_basis_column", "column": c} for c in pivot[rank:]]
    return {"type": "regression", "model": model}, warnings


def early_exercise_document(doc, report):
    """EarlyExerciseDiagnostics (mc_result.rs:249-273) of an American request
    from the device's pe_lsm_report (engine.LsmReport); the exercise-policy
    fingerprint is not carried."""
    lsm = doc["lsm"]
    degree = int(lsm["basis"]["max_degree"])
    columns = degree + 1
    models, diagnostics = [], []
    for m in report.decisions:
        model, warnings = _decision_document(m, columns, degree)
        models.append(model)
        diagnostics.append({"candidate_rows": int(report.training_trajectories), "feature_count": 1,
                            "itm_rows": int(m.itm_rows), "warnings": warnings})
    out = {
        "cpqr": {"abs_rank_tolerance": report.abs_rank_tolerance, "rel_rank_tolerance": report.rel_rank_tolerance},
        "decision_models": models,
        "dividend_collisions": list(report.dividend_collisions),
        "exercise_counts": [int(c) for c in report.exercise_counts],
        "exercise_dates": list(doc["product"]["exercise_dates"]),
        "exercise_probabilities": list(report.exercise_probabilities),
        "in_sample_value": report.in_sample_value,
        "itm_abs_tolerance": report.itm_abs_tolerance,
        "max_matrix_elements": int(report.max_matrix_elements),
        "policy_basis": {"exponents": [[k] for k in range(columns)], "feature_count": 1, "max_degree": degree},
        "regression_diagnostics": diagnostics,
        "training_random_domain": {"type": _DOMAINS[report.training_random_domain]},
        "training_sampling_units": int(report.training_sampling_units),
        "training_trajectories": int(report.training_trajectories),
        "valuation_random_domain": {"type": _DOMAINS[report.valuation_random_domain]},
        "valuation_sampling_units": int(report.valuation_sampling_units),
        "valuation_trajectories": int(report.valuation_trajectories),
    }
    if report.stopping_indices is not None:
        out["stopping_indices"] = [int(v) for v in report.stopping_indices]
    for k in ("training_direction_checksum", "training_scramble_checksum", "valuation_direction_checksum",
              "valuation_scramble_checksum"):
        if getattr(report, k) is not None:
            out[k] = getattr(report, k)
    return out


def result_document(doc, result, library_version="0.1.0", platform="mi355x-hip", report=None):
    """pricing_result document for the request document `doc` priced into the
    pe_result `result` (price and the Greeks the request asks for).  An
    American request with its pe_lsm_report (engine.LsmReport) also gets the
    monte_carlo block: the engine diagnostics and the early-exercise
    diagnostics."""
    engine = doc["engine"]["type"]
    estimator = "pseudo_monte_carlo" if engine == "pseudo_monte_carlo" else "randomized_quasi_monte_carlo"
    spot = float(doc["market"]["spot"])
    e = result.estimates

    def est(c, scale=1.0):
        # the value block keeps the engine's own 95% bracket
        if scale == 1.0:
            x = e[c]
            return {"value": x.value, "standard_error": x.standard_error, "confidence_lower": x.confidence_lower,
                    "confidence_upper": x.confidence_upper, "estimator": {"type": estimator},
                    "effective_sampling_units": int(x.effective_units)}
        return estimate_document(e[c].value, e[c].standard_error, e[c].effective_units, estimator, scale)

    risk = doc.get("risk", {})
    risks = {}
    if risk.get("delta"):
        risks["delta"] = {"raw": est(1), "raw_unit": {"type": "delta_raw"},
                          "market_scaled": est(1, spot * 0.01), "market_scaled_unit": {"type": "delta_one_percent_spot"}}
    if risk.get("gamma"):
        s = spot * 0.01
        risks["gamma"] = {"raw": est(3), "raw_unit": {"type": "gamma_raw"},
                          "market_scaled": est(3, s * s),
                          "market_scaled_unit": {"type": "gamma_one_percent_spot_squared"}}
    if risk.get("vega"):
        risks["vega"] = {"raw": est(2), "raw_unit": {"type": "vega_raw"},
                         "market_scaled": est(2, 0.01), "market_scaled_unit": {"type": "vega_one_vol_point"}}
    fingerprint = request_fingerprint(doc)
    out = {
        "document_kind": "pricing_result",
        "schema_version": 3,
        "value": est(0),
        "risks": risks,
        "diagnostics": {"warnings": []},
        "replay": {
            "schema_version": 3,
            "request_fingerprint": fingerprint,
            "library_version": library_version,
            "platform": platform,
            "migration": {"original_schema_version": 3, "current_schema_version": 3, "migration_ids": [],
                          "pre_migration_fingerprint": fingerprint, "post_migration_fingerprint": fingerprint},
        },
    }
    if doc["product"]["type"] == "american_vanilla" and report is not None:
        out["monte_carlo"] = {"diagnostics": monte_carlo_diagnostics_document(result.diagnostics),
                              "early_exercise": early_exercise_document(doc, report)}
    return out
