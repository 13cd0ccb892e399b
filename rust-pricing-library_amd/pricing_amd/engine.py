"""Python facade over the engine's C ABI, mirroring the reference's
``PricingPlan.compile(request, ...)`` / ``plan.evaluate()`` surface
(crates/pricing-python/src/lib.rs:347-395, crates/pricing/src/engine/risk/pricing.rs:53-88).

The HIP library is required: loading fails loudly when it is missing, and
there is no CPU fallback.
"""
import ctypes as C
import os

from . import abi

_LIB_PATH = os.environ.get("PE_ENGINE_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib",
                                                             "librpl_amd.so")
_lib = None

BACKENDS = {"fp64": abi.PE_BACKEND_GPU_FP64, "mixed": abi.PE_BACKEND_GPU_MIXED}


class EngineError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"pe error {status}: {message}")
        self.status = status


def library_path():
    return _LIB_PATH


def load():
    """Loads and binds librpl_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise FileNotFoundError(f"{_LIB_PATH} is missing; run __graft_entry__.build() or python -m pricing_amd.build")
    lib = C.CDLL(_LIB_PATH)
    abi.bind(lib)
    if lib.pe_abi_version() != abi.PE_ABI_VERSION:
        raise RuntimeError("ABI version mismatch")
    _lib = lib
    return lib


def _check(status):
    if status != abi.PE_OK:
        raise EngineError(status, load().pe_last_error().decode())


def init(device=0):
    _check(load().pe_init(device))


class PricingPlan:
    """A compiled plan (host FP64 compile + device tables)."""

    def __init__(self, handle, request):
        self._handle = C.c_void_p(handle)
        self._request = request  # keeps the request buffers alive

    @classmethod
    def compile(cls, request, policy=None):
        from . import request as _request

        lib = load()
        pol = policy if policy is not None else _request.exec_policy()
        out = C.c_void_p()
        _check(lib.pe_plan_compile(request.pointer(), C.byref(pol), C.byref(out)))
        return cls(out.value, request)

    def evaluate(self, backend="fp64"):
        res = abi.pe_result()
        _check(load().pe_plan_evaluate(self._handle, BACKENDS[backend], C.byref(res)))
        return res

    def evaluate_noalloc(self, backend="fp64"):
        """pe_plan_evaluate_noalloc: the same result with every device buffer
        in the library's fixed per-GPU workspace (BS / Black-76 plans)."""
        res = abi.pe_result()
        _check(load().pe_plan_evaluate_noalloc(self._handle, BACKENDS[backend], C.byref(res)))
        return res

    def free(self):
        if self._handle:
            load().pe_plan_free(self._handle)
            self._handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class RoughPlan(PricingPlan):
    """A lifted Heston plan on a Black-Scholes request carrier
    (RoughVolatilityPricingPlan::compile)."""

    @classmethod
    def compile(cls, carrier, model, maximum_step, policy=None):
        from . import request as _request

        lib = load()
        pol = policy if policy is not None else _request.exec_policy()
        out = C.c_void_p()
        _check(lib.pe_rough_plan_compile(carrier.pointer(), C.byref(model), maximum_step, C.byref(pol),
                                         C.byref(out)))
        plan = cls(out.value, carrier)
        plan._model = model
        return plan


def lifted_from_rough(hurst, factors, ratio):
    w = (C.c_double * factors)()
    x = (C.c_double * factors)()
    _check(load().pe_lifted_heston_from_rough(hurst, factors, ratio, w, x))
    n = 1 if hurst == 0.5 else factors
    return list(w)[:n], list(x)[:n]


def evaluate_shared(plans, backend="fp64"):
    """Evaluates plans sharing one simulated path set in one launch."""
    n = len(plans)
    handles = (C.c_void_p * n)(*[p._handle.value for p in plans])
    results = (abi.pe_result * n)()
    _check(load().pe_plans_evaluate_shared(handles, n, BACKENDS[backend], results))
    return list(results)


def evaluate_batch(plans, backend="fp64"):
    """Evaluates independent BS/Black-76 plans (a trade batch) in persistent launches."""
    n = len(plans)
    handles = (C.c_void_p * n)(*[p._handle.value for p in plans])
    results = (abi.pe_result * n)()
    _check(load().pe_batch_evaluate(handles, n, BACKENDS[backend], results))
    return list(results)


class LsmReport:
    """Early-exercise diagnostics of one plan: in-sample value, valuation
    exercise counts and probabilities, the decision model per non-terminal
    date, dividend collisions, optional stopping indices, the engines'
    domains, sizes and RQMC checksums, device ms (EarlyExerciseDiagnostics);
    for the network policy, its and the polynomial judge's out-of-sample
    values and whether the judge's policy was kept (the fallback)."""

    def __init__(self, r, counts, decisions, probabilities, collisions, stops):
        n = r.exercise_dates
        self.in_sample_value = r.in_sample_value
        self.training_ms = r.training_ms
        self.valuation_ms = r.valuation_ms
        self.path_ms = r.path_ms
        self.exercise_dates = n
        self.exercise_counts = list(counts)[:n]
        self.exercise_probabilities = list(probabilities)[:n]
        self.dividend_collisions = [bool(c) for c in list(collisions)[:n]]
        self.decisions = list(decisions)[:max(n - 1, 0)]
        self.stopping_indices = list(stops)[:r.valuation_trajectories] if stops is not None else None
        self.training_random_domain = r.training_random_domain
        self.valuation_random_domain = r.valuation_random_domain
        self.training_sampling_units = r.training_sampling_units
        self.training_trajectories = r.training_trajectories
        self.valuation_sampling_units = r.valuation_sampling_units
        self.valuation_trajectories = r.valuation_trajectories
        self.training_direction_checksum = bytes(r.training_direction_checksum).hex() if r.training_has_checksums else None
        self.training_scramble_checksum = bytes(r.training_scramble_checksum).hex() if r.training_has_checksums else None
        self.valuation_direction_checksum = (bytes(r.valuation_direction_checksum).hex()
                                             if r.valuation_has_checksums else None)
        self.valuation_scramble_checksum = (bytes(r.valuation_scramble_checksum).hex()
                                            if r.valuation_has_checksums else None)
        self.policy_feature_count = r.policy_feature_count
        self.policy_max_degree = r.policy_max_degree
        self.itm_abs_tolerance = r.itm_abs_tolerance
        self.abs_rank_tolerance = r.abs_rank_tolerance
        self.rel_rank_tolerance = r.rel_rank_tolerance
        self.max_matrix_elements = r.max_matrix_elements
        # network policy: its own and the polynomial judge's out-of-sample values
        self.network_price = abi.pe_estimate.from_buffer_copy(r.network_price)
        self.judge_price = abi.pe_estimate.from_buffer_copy(r.judge_price)
        self.judge_in_sample_value = r.judge_in_sample_value
        self.fallback_used = bool(r.fallback_used)
        self.network_training_ms = r.network_training_ms
        self.judge_training_ms = r.judge_training_ms
        self.network_valuation_ms = r.network_valuation_ms
        self.judge_valuation_ms = r.judge_valuation_ms
        self.network_rows_trained = r.network_rows_trained


def evaluate_lsm(plans, backend="fp64", max_dates=1024, stopping_capacity=0):
    """Prices American plans by least-squares Monte Carlo on the device;
    returns (results, reports).  With stopping_capacity > 0 each report also
    carries up to that many stopping indices (one per valuation trajectory)."""
    n = len(plans)
    handles = (C.c_void_p * n)(*[p._handle.value for p in plans])
    results = (abi.pe_result * n)()
    reports = (abi.pe_lsm_report * n)()
    keep = []
    for r in reports:
        counts = (C.c_uint64 * max_dates)()
        decisions = (abi.pe_lsm_decision * max_dates)()
        probabilities = (C.c_double * max_dates)()
        collisions = (C.c_int32 * max_dates)()
        stops = (C.c_uint32 * stopping_capacity)() if stopping_capacity else None
        r.capacity = max_dates
        r.exercise_counts = C.cast(counts, C.POINTER(C.c_uint64))
        r.decisions = C.cast(decisions, C.POINTER(abi.pe_lsm_decision))
        r.exercise_probabilities = C.cast(probabilities, C.POINTER(C.c_double))
        r.dividend_collisions = C.cast(collisions, C.POINTER(C.c_int32))
        if stops is not None:
            r.stopping_indices = C.cast(stops, C.POINTER(C.c_uint32))
            r.stopping_capacity = stopping_capacity
        keep.append((counts, decisions, probabilities, collisions, stops))
    _check(load().pe_lsm_evaluate(handles, n, BACKENDS[backend], results, reports))
    return list(results), [LsmReport(r, *k) for r, k in zip(reports, keep)]


def evaluate_lsm_dual(plan, continuation_inner_paths, exercise_inner_paths, inner_seed, backend="fp64"):
    """Andersen-Broadie lower/upper bounds of one American plan on the device
    (AndersenBroadiePlan::evaluate); returns the pe_dual_result."""
    out = abi.pe_dual_result()
    _check(load().pe_lsm_dual_evaluate(plan._handle, continuation_inner_paths, exercise_inner_paths, inner_seed,
                                       BACKENDS[backend], C.byref(out)))
    return out


def local_variance_from_ssvi(theta_times, theta_values, terminal_slope, rho, eta, gamma, time_nodes, x_nodes,
                             floor, cap):
    """Dupire local variance of a StandardSsvi surface, built in FP64 on the GPU."""
    lib = load()
    s = abi.pe_ssvi_surface()
    tt = (C.c_double * len(theta_times))(*theta_times)
    tv = (C.c_double * len(theta_values))(*theta_values)
    s.theta_count = len(theta_times)
    s.theta_times = C.cast(tt, C.POINTER(C.c_double))
    s.theta_values = C.cast(tv, C.POINTER(C.c_double))
    s.theta_terminal_slope = terminal_slope
    s.rho, s.eta, s.gamma = rho, eta, gamma
    t = (C.c_double * len(time_nodes))(*time_nodes)
    x = (C.c_double * len(x_nodes))(*x_nodes)
    out = (C.c_double * (len(time_nodes) * len(x_nodes)))()
    rep = C.c_uint32()
    _check(lib.pe_local_variance_from_ssvi(C.byref(s), len(time_nodes), t, len(x_nodes), x, floor, cap, out,
                                           C.byref(rep)))
    return list(out), rep.value


def last_kernel_ms():
    ms = C.c_double()
    n = C.c_uint32()
    load().pe_last_kernel_stats(C.byref(ms), C.byref(n))
    return ms.value, n.value
