"""Extracts golden vectors from the reference's own fixture files.

Run in a container where the reference is mounted (default /root/reference);
writes tests/golden/reference_vectors.json.  The output holds data only:
request documents, expected estimates and Decimal reference values copied
from

  fixtures/replay/{european_bs,local_volatility,path_dependence}-windows-x86_64.json
  fixtures/local-vol/reference-cases-v0.1.json  (standard_ssvi, constant_vol_dupire, affine_dividend)
  fixtures/path-dependence/reference-cases-v0.1.json  (smoothing, extrema, bridge, smoothed bridge, jumps)
  fixtures/rough-volatility/lifted-factor-prices.json  (finite lift kernels)
  fixtures/rough-volatility/two-step-prices.json  (lifted Heston two-step call prices)
  fixtures/replay/early_exercise-windows-x86_64.json  (LSM requests, prices, fitted decision models,
      early-exercise diagnostics)
  (replay cases also keep the execution's monte_carlo diagnostics block)
  fixtures/early-exercise/reference-cases-v0.1.json  (decision, basis, scaling, QR cases)
  fixtures/v3/*.golden.json  (schema-v3 request/result documents)
  fixtures/product-graph-v1|v2/*.json  (copied to product_graphs/)

The tests read the JSON this script writes; nothing at test time reads the
reference tree.
"""
import json
import os
import sys

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def replay(name):
    path = os.path.join(ROOT, "fixtures", "replay", name)
    with open(path) as f:
        doc = json.load(f)
    cases = []
    for case in doc["cases"]:
        result = case["result"]
        risks = {}
        for key, value in (result.get("risks") or {}).items():
            if isinstance(value, dict) and "raw" in value:
                risks[key] = {"value": value["raw"]["value"], "standard_error": value["raw"]["standard_error"]}
        cases.append({
            "name": case["name"],
            "source": f"fixtures/replay/{name}",
            "request": case["request"],
            "reduction_block_size": case["plan"]["reduction_block_size"],
            "worker_threads": case["plan"]["worker_threads"],
            "value": result["value"],
            "risks": risks,
            "sampling_variance_bits": case["execution"]["sampling_variance_bits"],
            "evaluated_paths": case["execution"]["evaluated_paths"],
            "risk_validation": case["execution"].get("risk_validation"),
            "barrier_bridge": (case.get("path_diagnostics") or {}).get("barrier_bridge"),
            "request_fingerprint": result["replay"]["request_fingerprint"],
            "result_risks": {k: v for k, v in (result.get("risks") or {}).items() if isinstance(v, dict) and "raw" in v},
            "monte_carlo": case["execution"]["monte_carlo"],
        })
    return cases


def main():
    out = {"replay": []}
    for name in ("european_bs-windows-x86_64.json", "local_volatility-windows-x86_64.json",
                 "path_dependence-windows-x86_64.json"):
        out["replay"].extend(replay(name))
    with open(os.path.join(ROOT, "fixtures", "local-vol", "reference-cases-v0.1.json")) as f:
        lv = json.load(f)
    out["standard_ssvi"] = [c for c in lv["standard_ssvi"] if "eta" in c["inputs"]]
    out["constant_vol_dupire"] = lv["constant_vol_dupire"]
    out["affine_dividend"] = lv["affine_dividend"]
    with open(os.path.join(ROOT, "fixtures", "path-dependence", "reference-cases-v0.1.json")) as f:
        pd = json.load(f)
    out["path_dependence"] = {k: pd[k] for k in ("smoothing_cases", "extrema_cases", "bridge_cases",
                                                "smoothed_bridge_cases", "jump_cases")}
    rough = os.path.join(ROOT, "fixtures", "rough-volatility")
    with open(os.path.join(rough, "lifted-factor-prices.json")) as f:
        lifted = json.load(f)
    with open(os.path.join(rough, "two-step-prices.json")) as f:
        two = json.load(f)
    out["lifted_kernels"] = lifted["kernels"]
    out["two_step_lifted_heston"] = {
        "spot": two["spot"], "time_nodes": two["time_nodes"], "strikes": two["strikes"],
        "case": next(c for c in two["cases"] if c["name"] == "lifted_heston")}
    with open(os.path.join(ROOT, "fixtures", "replay", "early_exercise-windows-x86_64.json")) as f:
        ee = json.load(f)
    out["early_exercise_replay"] = [{
        "name": c["name"],
        "request": c["request"],
        "reduction_block_size": c["plan"]["reduction_block_size"],
        "value": c["result"]["value"],
        "request_fingerprint": c["result"]["replay"]["request_fingerprint"],
        "risks": {k: {"value": v["raw"]["value"], "standard_error": v["raw"]["standard_error"]}
                  for k, v in (c["result"].get("risks") or {}).items() if isinstance(v, dict) and "raw" in v},
        "early_exercise": {k: v for k, v in c["result"]["monte_carlo"]["early_exercise"].items()
                           if k != "regression_diagnostics"},
        "monte_carlo": c["execution"]["monte_carlo"],
        "worker_threads": c["plan"]["worker_threads"],
    } for c in ee["cases"]]
    with open(os.path.join(ROOT, "fixtures", "early-exercise", "reference-cases-v0.1.json")) as f:
        out["early_exercise_cases"] = json.load(f)
    # schema-v3 golden documents (request and result wire round trips)
    out["v3_documents"] = {}
    for name in ("pricing_request.golden.json", "pricing_request_american.golden.json",
                 "pricing_result.golden.json", "pricing_result_american.golden.json"):
        with open(os.path.join(ROOT, "fixtures", "v3", name)) as f:
            out["v3_documents"][name] = json.load(f)
    # product_graph documents (fixtures/product-graph-v1|v2), copied as data
    os.makedirs(os.path.join(HERE, "product_graphs"), exist_ok=True)
    for sub, name in (("product-graph-v1", "coupon_and_capped_call.golden.json"), ("product-graph-v2", "digital.json")):
        with open(os.path.join(ROOT, "fixtures", sub, name)) as f:
            text = f.read()
        with open(os.path.join(HERE, "product_graphs", name), "w") as f:
            f.write(text)
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
