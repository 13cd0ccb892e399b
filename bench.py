#!/usr/bin/env python3
"""Headline benchmark: path-steps/s of the mixed-precision engine on the
local-vol Log-Euler path-dependent workload, next to the all-FP64 device mode
and the CPU oracle.

Workload (one "step" = one evaluation of the whole job):
  2^16 RQMC points x 16 scrambles = 2^20 paths, Sobol + LMS scrambles with
  Brownian-bridge ordering, Log-Euler local vol over the union of a 32 x 64
  Dupire grid (eSSVI-type StandardSsvi, theta(t) = 0.04 t, rho = -0.6,
  eta = 1.0, gamma = 0.4; built in FP64 on the GPU) and 252 daily fixings;
  products: up-and-out call K=100, B=130 with daily monitoring and an
  arithmetic Asian call K=100, priced on the same paths in one launch.

Further blocks, each with its own workload and roofline:
  * "european_bs": the single European call in the FP64 mode (price, Delta,
    Vega against the closed form);
  * "batch_bs": 4096 uneven BS/Black-76 trades with affine dividends and
    Delta/Vega, mixed vs FP64 device modes and the CPU oracle on a sample;
  * "rough_smile": the lifted Heston 21 x 8 European smile on one
    2^20-path RQMC set;
  * "rough_bermudan": a Bermudan put on the same model by least squares on
    (spot, variance);
  * "graph_payoff": product graphs through the device tape interpreter;
  * "lsm_american": 256 American puts by LSM with the weighted-Laguerre
    policy and with the learned (MLP) policy.

Prints one JSON line.  --gpus above 1 is refused: this build runs one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rust-pricing-library_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from pricing_amd import abi, engine, request, workloads  # noqa: E402

METRIC = "path-steps/s at matched RQMC std-error vs FP64 oracle; NN-LSM vs poly-LSM time"
FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak
FP64_PEAK_TFLOPS = 78.6    # MI355X published FP64 vector peak
HBM_PEAK_GBS = 8000.0      # MI355X published HBM3E bandwidth
# Frozen work model (flops per path-step of the local-vol path-dependent set):
# normal draw 34 + bridge 7 + LV Log-Euler step 24 + barrier monitor 4 +
# Asian accumulate 2 + observation 2.
FLOPS_PER_PATH_STEP = 73


def measured_traffic():
    """HBM bytes per launch of the mixed kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes, scaled by the same counters on a copy of
    known size (profiles/lv_kernel_traffic.json)."""
    path = os.path.join(ROOT, "profiles", "lv_kernel_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def lsm_traffic():
    """HBM bytes per launch of the stored-path LSM kernels from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes, scaled by the same counters on
    a copy of known size (profiles/lsm_kernel_traffic.json)."""
    try:
        with open(os.path.join(ROOT, "profiles", "lsm_kernel_traffic.json")) as f:
            return {k: v["traffic_bytes"] for k, v in json.load(f)["kernels"].items()}
    except (OSError, KeyError, ValueError):
        return {}


def build_plans(points, scrambles):
    expiry = workloads.weekday_dates(252)[-1] / 365.0
    grid = workloads.lv_dupire_grid(expiry, build=engine.local_variance_from_ssvi)
    reqs = workloads.lv_path_dependent(grid, points=points, scrambles=scrambles, rate=0.05)
    plans = [engine.PricingPlan.compile(r) for r in reqs]
    return reqs, plans, grid


def time_mode(plans, backend, steps, warmup):
    for _ in range(warmup):
        engine.evaluate_shared(plans, backend)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        res = engine.evaluate_shared(plans, backend)
        kernel_ms.append(engine.last_kernel_ms()[0])
    wall = time.perf_counter() - t0
    return res, wall, kernel_ms


def cpu_baseline(reqs, target_seconds, cores=None):
    """The CPU FP64 oracle on all host cores (or `cores` threads) on a bounded
    sample of the same workload: the same plans with fewer RQMC points and
    scrambles, grown until the sample takes about `target_seconds`."""
    from _support import oracle

    lib = oracle()
    cores = int(lib.oracle_hardware_threads()) if cores is None else cores
    points, scrambles = 1 << 10, 2
    while True:
        sample = []
        for r in reqs:
            q = request.Request()
            q.c = abi.pe_request.from_buffer_copy(r.c)
            q._keep = r._keep
            q.c.engine.units = points
            q.c.engine.scrambles = scrambles
            sample.append(q)
        arr = (C.POINTER(abi.pe_request) * len(sample))(*[q.pointer() for q in sample])
        out = (abi.pe_result * len(sample))()
        pol = request.exec_policy(cores, 64)
        t0 = time.perf_counter()
        rc = lib.oracle_evaluate_shared(arr, len(sample), C.byref(pol), out)
        dt = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(lib.oracle_last_error().decode())
        if dt >= 0.5 * target_seconds or scrambles >= 16:
            break
        grow = max(2, min(16, int(target_seconds / max(dt, 1e-3))))
        if points < 1 << 16:
            points = min(1 << 16, points * grow)
            points = 1 << (points.bit_length() - 1)
        else:
            scrambles = min(16, scrambles * grow)
    path_steps = out[0].path_steps
    return {"value": path_steps / dt, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "sample": f"{points} points x {scrambles} scrambles x {out[0].time_steps} steps ({path_steps} path-steps), "
                      f"barrier + Asian on shared paths, FP64 oracle, {dt:.2f} s wall"}


def european_block(no_cpu):
    """Config 1: European call under Black-Scholes in the FP64 device mode
    (S0 = K = 100, r = 5%, sigma = 20%, T = 1, one exact lognormal step,
    2^16 pseudo paths from seed 0x0123456789abcdef, pathwise Delta and Vega,
    block 4096), checked against the closed form.  At 2^16 paths one launch
    is latency-bound (a few microseconds of work), so the same plan is also
    timed at 2^24 paths for the FP64 kernel's roofline; work model: 120 FP64
    flops per path (normal draw, exact step, payoff, pathwise Delta/Vega and
    the three-component reductions)."""
    import math

    def closed_form():
        s, k, r, v, t = 100.0, 100.0, 0.05, 0.2, 1.0
        d1 = (math.log(s / k) + (r + 0.5 * v * v) * t) / (v * math.sqrt(t))
        d2 = d1 - v * math.sqrt(t)
        n = lambda x: 0.5 * math.erfc(-x / math.sqrt(2.0))
        return (s * n(d1) - k * math.exp(-r * t) * n(d2), n(d1),
                s * math.exp(-0.5 * d1 * d1) / math.sqrt(2.0 * math.pi) * math.sqrt(t))

    def timed(units, reps):
        plan = engine.PricingPlan.compile(workloads.european_bs(units=units),
                                          request.exec_policy(1, 4096))
        plan.evaluate("fp64")
        walls, kms = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            res = plan.evaluate("fp64")
            walls.append(time.perf_counter() - t0)
            kms.append(engine.last_kernel_ms()[0])
        return res, min(walls), sum(kms) / len(kms)

    price, delta, vega = closed_form()
    res, wall, kms = timed(1 << 16, 20)
    est = res.estimates
    block = {
        "workload": "European call, BS, S0=K=100, r=5%, q=0, sigma=20%, T=1, one exact step, 2^16 pseudo paths, "
                    "pathwise Delta/Vega, FP64 device mode",
        "paths": res.evaluated_paths,
        "wall_ms": 1e3 * wall, "kernel_ms": kms,
        "value": res.evaluated_paths / wall, "unit": "paths/s (one path-step each)",
        "price": {"mc": est[0].value, "se": est[0].standard_error, "closed_form": price,
                  "error_in_se": (est[0].value - price) / est[0].standard_error},
        "delta": {"mc": est[1].value, "se": est[1].standard_error, "closed_form": delta,
                  "error_in_se": (est[1].value - delta) / est[1].standard_error},
        "vega": {"mc": est[2].value, "se": est[2].standard_error, "closed_form": vega,
                 "error_in_se": (est[2].value - vega) / est[2].standard_error},
    }
    big, wall24, kms24 = timed(1 << 24, 5)
    flops = 120.0 * big.evaluated_paths
    block["throughput_2p24"] = {"paths": big.evaluated_paths, "wall_ms": 1e3 * wall24, "kernel_ms": kms24,
                                "value": big.evaluated_paths / wall24}
    block["roofline"] = {"bound": "valu", "kernel": "bs_batch_kernel<double, risk> at 2^24 paths",
                         "achieved": flops / (kms24 * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": flops / (kms24 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "flops_per_path": 120,
                         "traffic": None}
    if not no_cpu:
        from _support import oracle, oracle_evaluate

        cores = int(oracle().oracle_hardware_threads())
        req = workloads.european_bs(units=1 << 22)
        for n in (1, cores):
            t0 = time.perf_counter()
            r = oracle_evaluate(req, n, 4096)
            dt = time.perf_counter() - t0
            block["cpu_baseline" if n == cores else "cpu_baseline_one_core"] = {
                "value": r.evaluated_paths / dt, "unit": "paths/s", "cores": n, "kind": "port",
                "sample": f"the same plan at 2^22 paths, FP64 oracle, {dt:.2f} s wall"}
        block["speedup_vs_cpu_all_cores"] = block["throughput_2p24"]["value"] / block["cpu_baseline"]["value"]
    return block


def batch_block(steps, cpu_seconds, no_cpu, no_fp64):
    """Config 3: 4096 uneven BS/Black-76 trades with affine dividends, Delta and
    Vega, one launch per kernel flavour; mixed vs FP64 device modes and the CPU
    oracle on a sample of the trades."""
    reqs = workloads.batch_trades(4096)
    plans = [engine.PricingPlan.compile(r) for r in reqs]
    engine.evaluate_batch(plans, "mixed")
    t0 = time.perf_counter()
    for _ in range(steps):
        mixed = engine.evaluate_batch(plans, "mixed")
    wall = (time.perf_counter() - t0) / steps
    kernel_ms = engine.last_kernel_ms()[0]
    path_steps = sum(r.path_steps for r in mixed)
    block = {
        "workload": "4096 trades (European / cash Digital / discrete or continuous Barrier / Asian / Lookback), "
                    "K~U[70,130], sigma~U[0.1,0.5], T~U[0.1,5]y, 12-365 dates, 2^U[12,18] pseudo paths, "
                    "0-4 dividends (1.0 cash + 1%), compact-C2 half width 1.0, price + Delta + Vega",
        "path_steps": path_steps,
        "paths": sum(r.evaluated_paths for r in mixed),
        "value": path_steps / wall,
        "unit": "path-steps/s",
        "ms_per_step": 1e3 * wall,
        "kernel_ms": kernel_ms,
        # work model: normal draw 34 + BS step 8 + payoff/monitor 2-14 +
        # pathwise adjoint 8 = 64 flops per path-step for the risk trades
        "roofline": {"bound": "valu", "kernel": "bs_batch_kernel<float, *> (all flavours of the launch)",
                     "achieved": path_steps * 64 / (kernel_ms * 1e-3) / 1e12, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": path_steps * 64 / (kernel_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                     "flops_per_path_step": 64, "traffic": None},
    }
    if not no_fp64:
        engine.evaluate_batch(plans, "fp64")
        t0 = time.perf_counter()
        f64 = engine.evaluate_batch(plans, "fp64")
        wall64 = time.perf_counter() - t0
        block["fp64_value"] = path_steps / wall64
        block["fp64_kernel_ms"] = engine.last_kernel_ms()[0]
        block["mixed_over_fp64"] = wall64 / wall
        worst = {}
        for name, c in (("price", 0), ("delta", 1), ("vega", 2)):
            worst[name] = max(abs(a.estimates[c].value - b.estimates[c].value) / b.estimates[c].standard_error
                              for a, b in zip(mixed, f64) if b.estimates[c].standard_error > 0)
        block["max_mixed_minus_fp64_in_se"] = worst
    if not no_cpu:
        from concurrent.futures import ThreadPoolExecutor

        from _support import oracle_evaluate, oracle

        cores = int(oracle().oracle_hardware_threads())
        sample = [r for r in reqs if r.trade_id % 64 == 0]
        full_units = [r.c.engine.units for r in sample]
        # one trade per core at a time (the oracle releases the GIL); trades
        # longest-first so the pool drains evenly
        order = sorted(range(len(sample)), key=lambda i: -full_units[i] * max(1, sample[i].c.product.date_count))
        scale = 64
        with ThreadPoolExecutor(max_workers=cores) as pool:
            while True:
                for r, n in zip(sample, full_units):
                    r.c.engine.units = max(256, n // scale)
                t0 = time.perf_counter()
                cpu_steps = sum(pool.map(lambda i: oracle_evaluate(sample[i], 1, 4096).path_steps, order))
                dt = time.perf_counter() - t0
                if dt >= 0.25 * cpu_seconds or scale == 1:
                    break
                scale = max(1, scale // max(2, min(16, int(cpu_seconds / max(dt, 1e-3)))))
        block["cpu_baseline"] = {
            "value": cpu_steps / dt, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "sample": f"{len(sample)} of the 4096 trades (every 64th) at 1/{scale} of their path counts "
                      f"({cpu_steps} path-steps), FP64 oracle, one trade per core, {dt:.2f} s wall"}
        block["speedup_vs_cpu_all_cores"] = block["value"] / block["cpu_baseline"]["value"]
    return block


def graph_block(steps, no_fp64):
    """Graph-product interpreter: weekly arithmetic Asian calls (52 fixings,
    2^20 pseudo paths each) as product graphs (register tape in the BS batch
    kernel) and as the fixed Asian product on the same paths -- one trade
    (256 workgroups: one wave per SIMD, latency-bound) and a batch of 16
    trades (4096 workgroups); kernel time per mode."""
    block = {"workload": "weekly arithmetic Asian call, 52 fixings, 2^20 pseudo paths per trade, BS, price only"}
    for trades in (1, 16):
        for mode in ("mixed",) + (() if no_fp64 else ("fp64",)):
            for name, as_graph in (("graph", True), ("fixed", False)):
                plans = [engine.PricingPlan.compile(workloads.graph_asian(as_graph=as_graph, seed=0x6EA9_A51A + i))
                         for i in range(trades)]
                engine.evaluate_batch(plans, mode)
                t0 = time.perf_counter()
                for _ in range(steps):
                    rs = engine.evaluate_batch(plans, mode)
                wall = (time.perf_counter() - t0) / steps
                block[f"{name}_{mode}_x{trades}"] = {
                    "value": sum(r.path_steps for r in rs) / wall, "unit": "path-steps/s",
                    "kernel_ms": engine.last_kernel_ms()[0], "price": rs[0].estimates[0].value,
                    "se": rs[0].estimates[0].standard_error}
            block[f"graph_over_fixed_kernel_{mode}_x{trades}"] = (block[f"graph_{mode}_x{trades}"]["kernel_ms"] /
                                                                  block[f"fixed_{mode}_x{trades}"]["kernel_ms"])
    return block


def rough_block(steps, cpu_seconds, no_cpu, no_fp64):
    """Config 5 smile: lifted Heston (H=0.1, 20 geometric factors of ratio 2.5),
    21 strikes x 8 maturities priced on one 2^20-path RQMC set over the union
    grid of the maturities with maximum step 2/256."""
    model = workloads.lifted_heston_model(engine.lifted_from_rough(0.1, 20, 2.5))
    reqs = workloads.rough_smile()
    plans = [engine.RoughPlan.compile(r, model, 2.0 / 256) for r in reqs]
    engine.evaluate_shared(plans, "mixed")
    t0 = time.perf_counter()
    for _ in range(steps):
        mixed = engine.evaluate_shared(plans, "mixed")
    wall = (time.perf_counter() - t0) / steps
    path_steps = mixed[0].path_steps
    block = {
        "workload": "lifted Heston H=0.1 v0=0.02 kappa=0.3 theta=0.02 nu=0.3 rho=-0.7, 20 factors (ratio 2.5), "
                    "2^16 RQMC points x 16 scrambles, 21 strikes 0.8-1.2 S0 x 8 maturities 0.1-2y on one path set",
        "time_steps": mixed[0].time_steps,
        "paths": mixed[0].evaluated_paths,
        "path_steps": path_steps,
        "value": path_steps / wall,
        "unit": "path-steps/s",
        "ms_per_step": 1e3 * wall,
        "kernel_ms": engine.last_kernel_ms()[0],
        "atm_1y": {"price": mixed[4 * 21 + 10].estimates[0].value, "se": mixed[4 * 21 + 10].estimates[0].standard_error},
    }
    # work model: 168 flops per path-step (20-factor step with 2 normals +
    # payoff); the mixed scheme runs the variance process in FP64, so the
    # FP64 vector peak is the bound it is measured against
    kms = block["kernel_ms"]
    block["roofline"] = {"bound": "valu", "kernel": "rough_path_kernel<float, 20> + rough_payoff_kernel",
                         "achieved": path_steps * 168 / (kms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": path_steps * 168 / (kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                         "flops_per_path_step": 168, "traffic": None}
    if not no_fp64:
        engine.evaluate_shared(plans, "fp64")
        t0 = time.perf_counter()
        f64 = engine.evaluate_shared(plans, "fp64")
        wall64 = time.perf_counter() - t0
        block["fp64_value"] = path_steps / wall64
        block["fp64_kernel_ms"] = engine.last_kernel_ms()[0]
        block["mixed_over_fp64"] = wall64 / wall
        z = [(a.estimates[0].value - b.estimates[0].value) / b.estimates[0].standard_error
             for a, b in zip(mixed, f64) if b.estimates[0].standard_error > 0]
        block["max_mixed_minus_fp64_in_se"] = max(abs(v) for v in z)
        # the truncated scheme amplifies last-bit differences path by path:
        # the mean over the grid shows whether the gap is a bias or noise
        block["mean_mixed_minus_fp64_in_se"] = sum(z) / len(z)
    if not no_cpu:
        from _support import oracle, oracle_rough

        cores = int(oracle().oracle_hardware_threads())
        points = 1 << 8
        while True:
            sample = workloads.rough_smile(points=points, scrambles=16)
            t0 = time.perf_counter()
            res, _ = oracle_rough(sample, model, 2.0 / 256, workers=cores, block=64)
            dt = time.perf_counter() - t0
            if dt >= 0.25 * cpu_seconds or points >= 1 << 16:
                break
            points = min(1 << 16, points * max(2, min(16, int(cpu_seconds / max(dt, 1e-3)))))
            points = 1 << (points.bit_length() - 1)
        block["cpu_baseline"] = {
            "value": res[0].path_steps / dt, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "sample": f"the same 168-option smile on {points} points x 16 scrambles ({res[0].path_steps} path-steps), "
                      f"FP64 oracle, {dt:.2f} s wall"}
        block["speedup_vs_cpu_all_cores"] = block["value"] / block["cpu_baseline"]["value"]
    return block


def rough_bermudan_block(cpu_seconds, no_cpu, no_fp64):
    """Config 5 early exercise: a Bermudan put (12 dates, one year, K = S0 =
    100, r = 5%) on the smile model's lifted Heston paths, least squares on
    (spot, variance) with standardized monomials of total degree 2 and CPQR
    rules; 2^18 training + 2^18 independent valuation pseudo paths.  The
    policy trains on FP64 paths in both modes; the mixed mode values on the
    mixed scheme's paths.  The "network" leg prices the same Bermudan with the
    network continuation on (spot, variance)."""
    model = workloads.lifted_heston_model(engine.lifted_from_rough(0.1, 20, 2.5))
    req = workloads.rough_bermudan(units=1 << 18, train_units=1 << 18)
    plan = engine.RoughPlan.compile(req, model, 2.0 / 256)

    def run(backend, reps):
        engine.evaluate_lsm([plan], backend)
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            res, rep = engine.evaluate_lsm([plan], backend)
            wall = time.perf_counter() - t0
            if best is None or wall < best[2]:
                best = (res[0], rep[0], wall)
        return best

    mixed, mrep, mwall = run("mixed", 3)
    steps = mixed.time_steps
    path_steps = (req.c.lsm.training_engine.units + req.c.engine.units) * steps
    ms = mrep.training_ms + mrep.valuation_ms
    block = {
        "workload": "Bermudan put K=S0=100, r=5%, 12 exercise dates over 1y, lifted Heston H=0.1 (20 factors, "
                    "ratio 2.5), step 2/256; LSM on (spot, variance), total degree 2, CPQR; 2^18 training + "
                    "2^18 valuation pseudo paths",
        "time_steps": steps,
        "path_steps": path_steps,
        "device_ms": ms, "training_ms": mrep.training_ms, "path_ms": mrep.path_ms,
        "valuation_ms": mrep.valuation_ms, "wall_ms": 1e3 * mwall,
        "value": path_steps / (ms * 1e-3), "unit": "path-steps/s (training + valuation)",
        "price": mixed.estimates[0].value, "se": mixed.estimates[0].standard_error,
        "in_sample": mrep.in_sample_value,
        "early_exercise_share": 1.0 - mrep.exercise_counts[-1] / mrep.valuation_trajectories,
    }
    if not no_fp64:
        f64, frep, fwall = run("fp64", 1)
        fms = frep.training_ms + frep.valuation_ms
        block["fp64_device_ms"] = fms
        block["fp64_valuation_ms"] = frep.valuation_ms
        block["fp64_price"] = f64.estimates[0].value
        block["mixed_over_fp64"] = fms / ms
        block["mixed_minus_fp64_in_se"] = (mixed.estimates[0].value - f64.estimates[0].value) / \
            f64.estimates[0].standard_error
    # the network continuation on (spot, variance), judged by the polynomial
    # policy on the same paths (one launch prices both)
    nreq = workloads.rough_bermudan(units=1 << 18, train_units=1 << 18, method=abi.PE_LSM_MLP)
    nplan = engine.RoughPlan.compile(nreq, model, 2.0 / 256)
    engine.evaluate_lsm([nplan], "mixed")
    t0 = time.perf_counter()
    (nres,), (nrep,) = engine.evaluate_lsm([nplan], "mixed")
    nwall = time.perf_counter() - t0
    nn = {
        "config": "one 2-64-64-1 SiLU network per exercise date on (S/K, V), BF16 MFMA, Adam lr 1e-3, 4 epochs, "
                  "batch 8192, 2^18 training paths (in-the-money rows), control variate: Black value with the "
                  "expected average variance; the polynomial policy above is its judge on the same paths",
        "price": nrep.network_price.value, "se": nrep.network_price.standard_error,
        "minus_polynomial_in_se": (nrep.network_price.value - nrep.judge_price.value) / nrep.judge_price.standard_error,
        "network_training_ms": nrep.network_training_ms, "network_valuation_ms": nrep.network_valuation_ms,
        "time_over_polynomial": (nrep.network_training_ms + nrep.network_valuation_ms) / ms,
        "fallback_used": bool(nrep.fallback_used), "kept_price": nres.estimates[0].value,
        "rows_trained": int(nrep.network_rows_trained), "wall_ms": 1e3 * nwall,
    }
    if not no_fp64:
        (_,), (nf64,) = engine.evaluate_lsm([nplan], "fp64")
        nn["fp64_price"] = nf64.network_price.value
        nn["mixed_minus_fp64_in_se"] = (nrep.network_price.value - nf64.network_price.value) / \
            nf64.network_price.standard_error
    block["network"] = nn
    if not no_cpu:
        from _support import oracle_rough_lsm

        log2 = 10
        while True:
            sample = workloads.rough_bermudan(units=1 << log2, train_units=1 << log2)
            t0 = time.perf_counter()
            res, _, _, _, _ = oracle_rough_lsm(sample, model, 2.0 / 256)
            dt = time.perf_counter() - t0
            if dt >= 0.25 * cpu_seconds or log2 >= 18:
                break
            log2 = min(18, log2 + max(1, min(4, int(cpu_seconds / max(dt, 1e-3)).bit_length() - 1)))
        block["cpu_baseline"] = {
            "value": 2 * (1 << log2) * res.time_steps / dt, "unit": "path-steps/s (training + valuation)",
            "cores": 1, "kind": "port",
            "sample": f"the same Bermudan at 2^{log2} training + 2^{log2} valuation paths, FP64 oracle "
                      f"(serial path generation and regression), {dt:.2f} s wall"}
        block["speedup_vs_cpu_one_core"] = block["value"] / block["cpu_baseline"]["value"]
    return block


def lsm_block(cpu_seconds, no_cpu, no_fp64, no_mlp=False):
    """Config 4: 256 American puts by LSM, 50 exercise dates, 2^18 training +
    2^18 independent valuation paths per trade.  Plain policy: weighted
    Laguerre basis of degree 3 with FP64 Cholesky, fitted per trade and date.
    Learned policy: one 2-64-64-1 SiLU network per trade (inputs: moneyness
    and the Black value of the European option left), BF16 matrix cores,
    Adam lr 1e-3, 4 epochs, batch 8192, on the trade's 2^18 training paths,
    judged against the weighted-Laguerre policy on the same paths."""
    import numpy as np

    def run(method, train_log2, backend, reps=2):
        reqs = workloads.american_batch(256, train_log2=train_log2, value_log2=18, dates=50, method=method)
        plans = [engine.PricingPlan.compile(r) for r in reqs]
        engine.evaluate_lsm(plans, backend)
        walls = []
        for _ in range(reps):
            t0 = time.perf_counter()
            res, rep = engine.evaluate_lsm(plans, backend)
            walls.append(time.perf_counter() - t0)
        return reqs, res, rep, min(walls)

    reqs, poly, prep, pwall = run(abi.PE_LSM_POLY_LAGUERRE, 18, "mixed")
    # nominal path-steps: every training and valuation path over every date
    train_steps = sum(r.c.lsm.training_engine.units * r.c.product.date_count for r in reqs)
    value_steps = sum(r.c.engine.units * r.c.product.date_count for r in reqs)
    pms = prep[0].training_ms + prep[0].valuation_ms
    # Rooflines of the two stored-path stages (HBM-bound by design).  The
    # policy trains on FP64 paths in both modes.  The path kernel writes the
    # date-major FP64 training store once: 8 B per (trade, date, path).  The
    # induction (two launches per non-terminal date over every trade, then
    # the in-sample pass) has as algorithmic traffic the store read once (8 B
    # per path-date) plus the FP64 cash flow read and written once per
    # non-terminal date (16 B per path-date); its time is the HIP-event
    # interval from the end of the path kernel to the end of the induction.
    rep = prep[0]
    paths = sum(r.c.lsm.training_engine.units for r in reqs)
    dates = reqs[0].c.product.date_count
    induction_ms = rep.training_ms - rep.path_ms
    induction_bytes = paths * (8 * dates + 16 * (dates - 1))
    store_bytes = paths * dates * 8
    traffic = lsm_traffic()
    block = {
        "workload": "256 American puts, S0=100, K~U[80,120], sigma~U[0.2,0.4], r=3%, T=1y, 50 exercise dates, "
                    "2^18 training + 2^18 valuation pseudo paths per trade",
        "roofline": {
            "bound": "hbm", "kernel": "lsm_pass_kernel<double> + lsm_fit_kernel<double> per date (induction)",
            "avg_kernel_ms": induction_ms,
            "algorithmic_bytes": induction_bytes,
            "achieved": induction_bytes / (induction_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": induction_bytes / (induction_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": traffic.get("lsm_induction<double>"),
        },
        "path_store_roofline": {
            "bound": "hbm", "kernel": "lsm_train_paths_kernel<double>", "avg_kernel_ms": rep.path_ms,
            "algorithmic_bytes": store_bytes,
            "achieved": store_bytes / (rep.path_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": store_bytes / (rep.path_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": traffic.get("lsm_train_paths_kernel<double>"),
        },
        "poly_laguerre": {
            "device_ms": pms, "training_ms": prep[0].training_ms, "valuation_ms": prep[0].valuation_ms,
            "wall_ms": 1e3 * pwall,
            "path_steps": train_steps + value_steps,
            "value": (train_steps + value_steps) / (pms * 1e-3), "unit": "path-steps/s (training + valuation)",
            "mean_price": float(np.mean([r.estimates[0].value for r in poly])),
            "mean_se": float(np.mean([r.estimates[0].standard_error for r in poly])),
        },
    }
    if not no_fp64:
        _, f64, frep, fwall = run(abi.PE_LSM_POLY_LAGUERRE, 18, "fp64", reps=1)
        fms = frep[0].training_ms + frep[0].valuation_ms
        block["poly_laguerre"]["fp64_device_ms"] = fms
        block["poly_laguerre"]["mixed_over_fp64"] = fms / pms
        block["poly_laguerre"]["max_mixed_minus_fp64_in_se"] = max(
            abs(a.estimates[0].value - b.estimates[0].value) / b.estimates[0].standard_error for a, b in zip(poly, f64))
    if not no_cpu:
        from concurrent.futures import ThreadPoolExecutor

        from _support import oracle, oracle_lsm

        cores = int(oracle().oracle_hardware_threads())
        log2 = 10
        while True:
            sample = workloads.american_batch(cores, train_log2=log2, value_log2=log2, dates=50)
            t0 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=cores) as pool:
                out = list(pool.map(lambda r: oracle_lsm(r, 4096, 1)[0], sample))
            dt = time.perf_counter() - t0
            if dt >= 0.25 * cpu_seconds or log2 >= 18:
                break
            log2 += max(1, min(4, int(np.log2(max(2.0, cpu_seconds / max(dt, 1e-3))))))
            log2 = min(log2, 18)
        steps = len(out) * 2 * (1 << log2) * 50
        block["cpu_baseline"] = {
            "value": steps / dt, "unit": "path-steps/s (training + valuation)", "cores": cores, "kind": "port",
            "sample": f"{cores} trades of the same workload at 2^{log2} training + 2^{log2} valuation paths, "
                      f"weighted-Laguerre LSM in the FP64 oracle, one trade per core, {dt:.2f} s wall"}
        block["speedup_vs_cpu_all_cores"] = block["poly_laguerre"]["value"] / block["cpu_baseline"]["value"]
    if no_mlp:
        return block
    # learned policy at the stated configuration: 2^18 training paths per
    # trade, 4 epochs, batch 8192, Adam lr 1e-3, one network per trade; the
    # weighted-Laguerre judge runs in the same launch on the same paths
    _, nn, nrep, nwall = run(abi.PE_LSM_MLP, 18, "mixed", reps=1)
    nms = nrep[0].network_training_ms + nrep[0].network_valuation_ms
    se = np.array([r.judge_price.standard_error for r in nrep])
    z_judge = np.array([(r.network_price.value - r.judge_price.value) for r in nrep]) / se
    z_poly = np.array([r.network_price.value - b.estimates[0].value for r, b in zip(nrep, poly)]) / \
        np.array([b.estimates[0].standard_error for b in poly])
    rows = sum(r.network_rows_trained for r in nrep)
    kernel_ms = nrep[0].network_training_ms - nrep[0].path_ms
    # algorithmic flops per row: forward 2x64 + 64x64 + 64 = 4288 MACs, the
    # backward pass twice that, 2 flops per MAC
    flops = rows * 3 * 4288 * 2
    block["mlp"] = {
        "config": "one 2-64-64-1 SiLU network per trade (BF16 MFMA, FP32 accumulation), Adam lr 1e-3, 4 epochs, "
                  "batch 8192, 2^18 training paths per trade (in-the-money rows), 50 dates; weighted-Laguerre "
                  "judge and fallback on the same training and valuation paths",
        "training_paths_per_trade": 1 << 18,
        "network_training_ms": nrep[0].network_training_ms, "network_valuation_ms": nrep[0].network_valuation_ms,
        "device_ms": nms, "wall_ms": 1e3 * nwall,
        "judge_training_ms": nrep[0].judge_training_ms, "judge_valuation_ms": nrep[0].judge_valuation_ms,
        "time_over_poly": nms / pms,
        "mean_price": float(np.mean([r.network_price.value for r in nrep])),
        "mean_minus_judge_in_se": float(np.mean(z_judge)),
        "max_abs_minus_judge_in_se": float(np.max(np.abs(z_judge))),
        "mean_minus_poly_in_se": float(np.mean(z_poly)),
        "max_abs_minus_poly_in_se": float(np.max(np.abs(z_poly))),
        "fallbacks": int(sum(r.fallback_used for r in nrep)),
        "rows_trained": int(rows),
        "mfma": {"kernel": "lsm_mlp_kernel<float>", "kernel_ms": kernel_ms, "algorithmic_tflops": flops / 1e12,
                 "achieved": flops / (kernel_ms * 1e-3) / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                 "frac": flops / (kernel_ms * 1e-3) / 1e12 / 2500.0},
    }
    return block


def copy_probe(nbytes):
    """A plain device-to-device copy of known size: the HBM counters it reads
    (FETCH_SIZE, WRITE_SIZE) calibrate the per-dispatch traffic figures."""
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_bytes": nbytes}))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=1 << 16)
    ap.add_argument("--scrambles", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-european", action="store_true", help="skip the single European (FP64) block")
    ap.add_argument("--no-batch", action="store_true", help="skip the 4096-trade BS/Black-76 batch block")
    ap.add_argument("--no-rough", action="store_true", help="skip the lifted Heston smile block")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-product interpreter block")
    ap.add_argument("--no-lsm", action="store_true", help="skip the 256-trade American LSM block")
    ap.add_argument("--no-mlp", action="store_true",
                    help="skip the learned-policy leg of the LSM block (its cooperative launch is not profiled)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        return 2

    engine.init(0)
    reqs, plans, grid = build_plans(args.points, args.scrambles)
    res, wall, kms = time_mode(plans, "mixed", args.steps, args.warmup)
    path_steps = res[0].path_steps
    ms_per_step = 1e3 * wall / args.steps
    value = path_steps / (wall / args.steps)
    avg_kernel_ms = sum(kms) / len(kms)
    achieved = path_steps * FLOPS_PER_PATH_STEP / (avg_kernel_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "local-vol Log-Euler path-dependent set: 2^16 RQMC points x 16 scrambles (2^20 paths), "
                        f"{res[0].time_steps} steps (252 daily fixings + 32x64 Dupire grid nodes), Sobol+LMS, "
                        "Brownian bridge, up-and-out call K=100 B=130 + arithmetic Asian call K=100",
            "paths": res[0].evaluated_paths,
            "time_steps": res[0].time_steps,
            "path_steps_per_step": path_steps,
            "precision": "mixed: FP32 inverse normal, exp/sqrt, variance lookup and Brownian bridge; FP64 "
                         "log-state, payoff accumulators and statistics",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": measured_traffic(),
            "traffic_source": "committed calibration: rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel scaled by "
                              "the same counters on a copy of known size (profiles/lv_kernel_traffic.json)",
            "kernel": "lv_kernel<float>",
            "avg_kernel_ms": avg_kernel_ms,
            "flops_per_path_step": FLOPS_PER_PATH_STEP,
        },
        "results": {
            "barrier": {"price": res[0].estimates[0].value, "se": res[0].estimates[0].standard_error},
            "asian": {"price": res[1].estimates[0].value, "se": res[1].estimates[0].standard_error},
            "dupire_repairs": grid[5],
        },
    }
    if not args.no_fp64:
        f64, wall64, kms64 = time_mode(plans, "fp64", max(2, args.steps // 2), 1)
        v64 = path_steps / (wall64 / max(2, args.steps // 2))
        out["fp64_mode"] = {
            "value": v64,
            "ms_per_step": 1e3 * wall64 / max(2, args.steps // 2),
            "avg_kernel_ms": sum(kms64) / len(kms64),
            "mixed_over_fp64": value / v64,
            "barrier": {"price": f64[0].estimates[0].value, "se": f64[0].estimates[0].standard_error},
            "asian": {"price": f64[1].estimates[0].value, "se": f64[1].estimates[0].standard_error},
            "mixed_minus_fp64_in_se": [
                (res[i].estimates[0].value - f64[i].estimates[0].value) / f64[i].estimates[0].standard_error
                for i in range(2)],
            "se_ratio_mixed_over_fp64": [res[i].estimates[0].standard_error / f64[i].estimates[0].standard_error
                                         for i in range(2)],
        }
    if not args.no_cpu:
        cb = cpu_baseline(reqs, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu_all_cores"] = value / cb["value"]
        one = cpu_baseline(reqs, 0.5 * args.cpu_seconds, cores=1)
        out["cpu_baseline_one_core"] = one
        out["speedup_vs_cpu_one_core"] = value / one["value"]
    else:
        out["cpu_baseline"] = None
    if not args.no_european:
        out["european_bs"] = european_block(args.no_cpu)
    if not args.no_batch:
        out["batch_bs"] = batch_block(3, args.cpu_seconds, args.no_cpu, args.no_fp64)
    if not args.no_rough:
        out["rough_smile"] = rough_block(5, args.cpu_seconds, args.no_cpu, args.no_fp64)
        out["rough_bermudan"] = rough_bermudan_block(args.cpu_seconds, args.no_cpu, args.no_fp64)
    if not args.no_graph:
        out["graph_payoff"] = graph_block(5, args.no_fp64)
    if not args.no_lsm:
        out["lsm_american"] = lsm_block(args.cpu_seconds, args.no_cpu, args.no_fp64, args.no_mlp)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
