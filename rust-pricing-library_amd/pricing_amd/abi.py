"""ctypes mirror of include/pe.h (the engine's C ABI)."""
import ctypes as C

PE_ABI_VERSION = 3

# pe_status
PE_OK = 0
PE_ERR_INVALID_ARGUMENT = 1
PE_ERR_UNSUPPORTED = 2
PE_ERR_NONFINITE_STATE = 3
PE_ERR_DEVICE = 4
PE_ERR_ALLOCATION = 5
PE_ERR_INSUFFICIENT_UNITS = 6
PE_ERR_SIZE_LIMIT = 7

# pe_backend
PE_BACKEND_GPU_FP64 = 1
PE_BACKEND_GPU_MIXED = 2

# pe_model_kind
PE_MODEL_BLACK_SCHOLES = 0
PE_MODEL_BLACK76 = 1
PE_MODEL_LOCAL_VOL = 2

# pe_product_kind
PE_PRODUCT_EUROPEAN = 0
PE_PRODUCT_DIGITAL = 1
PE_PRODUCT_BARRIER = 2
PE_PRODUCT_ASIAN = 3
PE_PRODUCT_LOOKBACK = 4
PE_PRODUCT_AMERICAN = 5
PE_PRODUCT_GRAPH = 6

# pe_graph_opcode, by the JSON operation names of product_graph documents
GRAPH_OPCODES = {
    "literal": 0, "terminal_spot": 1, "pre_dividend_spot": 2, "add": 3, "subtract": 4, "multiply": 5,
    "divide": 6, "minimum": 7, "maximum": 8, "indicator": 9, "negate": 10, "barrier_hit_state": 11,
    "smooth_minimum": 12, "smooth_maximum": 13, "smooth_indicator": 14, "log": 15, "exp": 16,
    "select": 17, "compare": 18,
}
GRAPH_COMPARISONS = {"less": 0, "less_equal": 1, "equal": 2, "not_equal": 3, "greater_equal": 4, "greater": 5}

PE_CALL = 0
PE_PUT = 1

PE_ENGINE_PSEUDO = 0
PE_ENGINE_RQMC = 1

COMPONENTS = 18
COMPONENT_NAMES = (
    "price", "delta", "vega", "gamma", "bump_delta", "delta_difference", "bump_vega",
    "vega_difference", "bump_gamma", "gamma_difference", "barrier_endpoint_hit",
    "barrier_dividend_jump_hit", "barrier_bridge_hit_weight", "barrier_interval_count",
    "barrier_finite_correction_count", "barrier_zero_variance_count",
    "barrier_survival_underflow_count", "barrier_certain_survival_count",
)
# pe_component indices (PE_C_PRICE = 0 ... PE_C_BARRIER_CERTAIN_SURVIVAL_COUNT = 17)
for _i, _n in enumerate(COMPONENT_NAMES):
    globals()["PE_C_" + _n.upper()] = _i

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


class pe_curve(C.Structure):
    _fields_ = [("count", C.c_uint32), ("times", _DP), ("discount_factors", _DP)]


class pe_dividend_quote(C.Structure):
    _fields_ = [("ex_time", C.c_double), ("fixed_cash", C.c_double), ("beta", C.c_double)]


class pe_market(C.Structure):
    _fields_ = [
        ("spot", C.c_double),
        ("discount_curve", pe_curve),
        ("dividend_curve", pe_curve),
        ("dividend_count", C.c_uint32),
        ("dividends", C.POINTER(pe_dividend_quote)),
    ]


class pe_local_variance_grid(C.Structure):
    _fields_ = [
        ("time_count", C.c_uint32),
        ("x_count", C.c_uint32),
        ("time_nodes", _DP),
        ("x_nodes", _DP),
        ("values", _DP),
        ("floor", C.c_double),
        ("cap", C.c_double),
    ]


class pe_model(C.Structure):
    _fields_ = [("kind", C.c_int), ("volatility", C.c_double), ("lv", pe_local_variance_grid)]


class pe_graph_node(C.Structure):
    _fields_ = [
        ("id", C.c_uint32),
        ("opcode", C.c_int),
        ("operands", C.c_uint32 * 3),
        ("comparison", C.c_int),
        ("underlying_id", C.c_uint32),
        ("observation_day", C.c_int32),
        ("value", C.c_double),
    ]


class pe_graph_cashflow(C.Structure):
    _fields_ = [("output_node", C.c_uint32), ("payment_day", C.c_int32)]


class pe_product(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("side", C.c_int),
        ("expiry_day", C.c_int32),
        ("payment_day", C.c_int32),
        ("strike", C.c_double),
        ("notional", C.c_double),
        ("payout", C.c_double),
        ("payout_is_asset", C.c_int32),
        ("barrier", C.c_double),
        ("direction_down", C.c_int32),
        ("knock_in", C.c_int32),
        ("continuous", C.c_int32),
        ("has_rebate", C.c_int32),
        ("rebate", C.c_double),
        ("historical_hit", C.c_int32),
        ("has_historical_extremum", C.c_int32),
        ("historical_extremum", C.c_double),
        ("date_count", C.c_uint32),
        ("dates", _IP),
        ("weights", _DP),
        ("fixings", _DP),
        ("graph_node_count", C.c_uint32),
        ("graph_nodes", C.POINTER(pe_graph_node)),
        ("graph_cashflow_count", C.c_uint32),
        ("graph_cashflows", C.POINTER(pe_graph_cashflow)),
        ("graph_underlying_count", C.c_uint32),
        ("graph_underlyings", C.POINTER(C.c_uint32)),
    ]


class pe_engine(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("seed", C.c_uint64),
        ("units", C.c_uint64),
        ("scrambles", C.c_uint32),
        ("antithetic", C.c_int32),
        ("brownian_bridge", C.c_int32),
    ]


class pe_risk(C.Structure):
    _fields_ = [
        ("delta", C.c_int32),
        ("vega", C.c_int32),
        ("gamma", C.c_int32),
        ("gamma_relative", C.c_int32),
        ("gamma_bump", C.c_double),
        ("smoothing", C.c_int32),
        ("smoothing_half_width", C.c_double),
        ("aad_tile_capacity", C.c_uint32),
        ("checkpoint_interval", C.c_uint32),
    ]


PE_LSM_POLY_CPQR = 0
PE_LSM_POLY_LAGUERRE = 1
PE_LSM_MLP = 2


class pe_lsm(C.Structure):
    _fields_ = [
        ("training_engine", pe_engine),
        ("max_degree", C.c_uint32),
        ("abs_rank_tolerance", C.c_double),
        ("rel_rank_tolerance", C.c_double),
        ("itm_abs_tolerance", C.c_double),
        ("max_matrix_elements", C.c_uint64),
        ("method", C.c_int),
        ("mlp_hidden", C.c_uint32),
        ("mlp_epochs", C.c_uint32),
        ("mlp_batch", C.c_uint32),
        ("mlp_learning_rate", C.c_double),
        ("mlp_seed", C.c_uint64),
    ]


class pe_request(C.Structure):
    _fields_ = [
        ("market", pe_market),
        ("model", pe_model),
        ("product", pe_product),
        ("engine", pe_engine),
        ("risk", pe_risk),
        ("lsm", pe_lsm),
    ]


class pe_exec_policy(C.Structure):
    _fields_ = [("worker_threads", C.c_uint32), ("reduction_block_size", C.c_uint64)]


class pe_estimate(C.Structure):
    _fields_ = [
        ("value", C.c_double),
        ("standard_error", C.c_double),
        ("confidence_lower", C.c_double),
        ("confidence_upper", C.c_double),
        ("effective_units", C.c_uint64),
    ]


PE_ESTIMATOR_PSEUDO_MONTE_CARLO = 0
PE_ESTIMATOR_RQMC = 1
CURVE_REGIONS = ("pillar", "interpolated", "right_extrapolated")
VALUATION_KINDS = ("exact_contractual", "smoothed_surrogate")
PE_DOMAIN_VALUATION = 0
PE_DOMAIN_LSM_TRAIN = 1
PE_DOMAIN_RQMC_SCRAMBLE = 2
RANDOM_DOMAINS = ("valuation", "lsm_train", "rqmc_scramble")


class pe_mc_diagnostics(C.Structure):
    _fields_ = [
        ("master_seed", C.c_uint64),
        ("estimator", C.c_int),
        ("scramble_count", C.c_uint32),
        ("has_checksums", C.c_int32),
        ("direction_checksum", C.c_uint8 * 32),
        ("scramble_checksum", C.c_uint8 * 32),
        ("policy_version", C.c_uint32),
        ("worker_threads", C.c_uint32),
        ("reduction_block_size", C.c_uint64),
        ("aad_tile_policy_version", C.c_uint32),
        ("aad_tile_capacity", C.c_uint32),
        ("checkpoint_policy_version", C.c_uint32),
        ("checkpoint_interval", C.c_uint32),
        ("antithetic", C.c_int32),
        ("discount_region", C.c_int),
        ("dividend_region", C.c_int),
        ("valuation_kind", C.c_int),
    ]


class pe_result(C.Structure):
    _fields_ = [
        ("estimates", pe_estimate * COMPONENTS),
        ("component_sum", C.c_double * COMPONENTS),
        ("component_mean", C.c_double * COMPONENTS),
        ("component_m2", C.c_double * COMPONENTS),
        ("sampling_variance", C.c_double),
        ("estimator_variance", C.c_double),
        ("independent_units", C.c_uint64),
        ("evaluated_paths", C.c_uint64),
        ("path_steps", C.c_uint64),
        ("time_steps", C.c_uint32),
        ("sobol_dimension", C.c_uint32),
        ("device_ms", C.c_double),
        ("failing_unit", C.c_uint64),
        ("diagnostics", pe_mc_diagnostics),
    ]


class pe_ssvi_surface(C.Structure):
    _fields_ = [
        ("theta_count", C.c_uint32),
        ("theta_times", _DP),
        ("theta_values", _DP),
        ("theta_terminal_slope", C.c_double),
        ("rho", C.c_double),
        ("eta", C.c_double),
        ("gamma", C.c_double),
    ]


# Every symbol include/pe.h declares, with (restype, argtypes).
class pe_lifted_heston(C.Structure):
    _fields_ = [
        ("initial_variance", C.c_double),
        ("mean_reversion", C.c_double),
        ("long_run_variance", C.c_double),
        ("vol_of_vol", C.c_double),
        ("correlation", C.c_double),
        ("factor_count", C.c_uint32),
        ("weights", _DP),
        ("rates", _DP),
    ]


def lifted_heston(v0, kappa, theta, nu, rho, weights, rates):
    """A pe_lifted_heston with its arrays kept alive on the struct."""
    m = pe_lifted_heston()
    w = (C.c_double * len(weights))(*weights)
    x = (C.c_double * len(rates))(*rates)
    m.initial_variance, m.mean_reversion, m.long_run_variance = v0, kappa, theta
    m.vol_of_vol, m.correlation = nu, rho
    m.factor_count = len(weights)
    m.weights = C.cast(w, _DP)
    m.rates = C.cast(x, _DP)
    m._keep = (w, x)
    return m


class pe_lsm_decision(C.Structure):
    _fields_ = [
        ("continue_all", C.c_int32),
        ("rank", C.c_uint32),
        ("itm_rows", C.c_uint64),
        ("coefficients", C.c_double * 6),
        ("diagonal_abs", C.c_double * 6),
        ("pivot_order", C.c_uint32 * 6),
        ("feature_mean", C.c_double),
        ("feature_population_variance", C.c_double),
        ("feature_scale", C.c_double),
        ("feature_inactive", C.c_int32),
        ("active_columns", C.c_uint32),
        ("rank_threshold", C.c_double),
        ("residual_sum_squares", C.c_double),
        ("region_left", C.c_int32),
        ("region_cut_count", C.c_uint32),
        ("region_cuts", C.c_float * 64),
        ("zero_scale_threshold", C.c_double),
        ("variance_mean", C.c_double),
        ("variance_population_variance", C.c_double),
        ("variance_scale", C.c_double),
        ("variance_zero_scale_threshold", C.c_double),
        ("variance_inactive", C.c_int32),
        ("feature_count", C.c_uint32),
    ]


class pe_dual_result(C.Structure):
    _fields_ = [
        ("lower_bound", pe_estimate),
        ("upper_bound", pe_estimate),
        ("duality_gap", pe_estimate),
        ("price_confidence_interval_95", C.c_double * 2),
        ("outer_trajectories", C.c_uint64),
        ("exercise_date_count", C.c_uint32),
        ("pad", C.c_uint32),
        ("device_ms", C.c_double),
    ]


class pe_lsm_report(C.Structure):
    _fields_ = [
        ("in_sample_value", C.c_double),
        ("training_ms", C.c_double),
        ("valuation_ms", C.c_double),
        ("exercise_dates", C.c_uint32),
        ("capacity", C.c_uint32),
        ("exercise_counts", C.POINTER(C.c_uint64)),
        ("decisions", C.POINTER(pe_lsm_decision)),
        ("path_ms", C.c_double),
        ("exercise_probabilities", _DP),
        ("dividend_collisions", C.POINTER(C.c_int32)),
        ("stopping_indices", C.POINTER(C.c_uint32)),
        ("stopping_capacity", C.c_uint64),
        ("training_random_domain", C.c_int),
        ("valuation_random_domain", C.c_int),
        ("training_sampling_units", C.c_uint64),
        ("training_trajectories", C.c_uint64),
        ("valuation_sampling_units", C.c_uint64),
        ("valuation_trajectories", C.c_uint64),
        ("training_has_checksums", C.c_int32),
        ("valuation_has_checksums", C.c_int32),
        ("training_direction_checksum", C.c_uint8 * 32),
        ("training_scramble_checksum", C.c_uint8 * 32),
        ("valuation_direction_checksum", C.c_uint8 * 32),
        ("valuation_scramble_checksum", C.c_uint8 * 32),
        ("policy_feature_count", C.c_uint32),
        ("policy_max_degree", C.c_uint32),
        ("itm_abs_tolerance", C.c_double),
        ("abs_rank_tolerance", C.c_double),
        ("rel_rank_tolerance", C.c_double),
        ("max_matrix_elements", C.c_uint64),
        ("network_price", pe_estimate),
        ("judge_price", pe_estimate),
        ("judge_in_sample_value", C.c_double),
        ("fallback_used", C.c_int32),
        ("pad", C.c_uint32),
        ("network_training_ms", C.c_double),
        ("judge_training_ms", C.c_double),
        ("network_valuation_ms", C.c_double),
        ("judge_valuation_ms", C.c_double),
        ("network_rows_trained", C.c_uint64),
    ]


PE_SYMBOLS = {
    "pe_init": (C.c_int, [C.c_int32]),
    "pe_last_error": (C.c_char_p, []),
    "pe_abi_version": (C.c_int32, []),
    "pe_plan_compile": (C.c_int, [C.POINTER(pe_request), C.POINTER(pe_exec_policy), C.POINTER(C.c_void_p)]),
    "pe_plan_free": (None, [C.c_void_p]),
    "pe_plan_evaluate": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(pe_result)]),
    "pe_plan_evaluate_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(pe_result)]),
    "pe_plans_evaluate_shared": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_int, C.POINTER(pe_result)]),
    "pe_batch_evaluate": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_int, C.POINTER(pe_result)]),
    "pe_lifted_heston_from_rough": (C.c_int, [C.c_double, C.c_uint32, C.c_double, _DP, _DP]),
    "pe_rough_plan_compile": (C.c_int, [C.POINTER(pe_request), C.POINTER(pe_lifted_heston), C.c_double,
                                        C.POINTER(pe_exec_policy), C.POINTER(C.c_void_p)]),
    "pe_local_variance_from_ssvi": (
        C.c_int,
        [C.POINTER(pe_ssvi_surface), C.c_uint32, _DP, C.c_uint32, _DP, C.c_double, C.c_double, _DP,
         C.POINTER(C.c_uint32)],
    ),
    "pe_lsm_evaluate": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_int, C.POINTER(pe_result),
                                  C.POINTER(pe_lsm_report)]),
    "pe_lsm_dual_evaluate": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int,
                                       C.POINTER(pe_dual_result)]),
    "pe_mlp_gradients": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32,
                                   C.POINTER(C.c_float), _DP]),
    "pe_mlp_forward": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_float)]),
    "pe_last_kernel_stats": (C.c_int, [_DP, C.POINTER(C.c_uint32)]),
    "pe_plan_diagnostics": (C.c_int, [C.c_void_p, C.POINTER(pe_mc_diagnostics)]),
    "pe_normal_mixed_max_error": (C.c_int, [C.c_uint64, C.c_uint64, _DP, C.POINTER(C.c_uint32)]),
    "pe_blake3_256": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_uint8)]),
}


def bind(lib, symbols=PE_SYMBOLS):
    for name, (restype, argtypes) in symbols.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def estimate_dict(e):
    return {
        "value": e.value,
        "standard_error": e.standard_error,
        "confidence_lower": e.confidence_lower,
        "confidence_upper": e.confidence_upper,
        "effective_sampling_units": e.effective_units,
    }


def result_dict(r):
    out = {name: estimate_dict(r.estimates[i]) for i, name in enumerate(COMPONENT_NAMES)}
    out["sampling_variance"] = r.sampling_variance
    out["estimator_variance"] = r.estimator_variance
    out["independent_sampling_units"] = r.independent_units
    out["evaluated_paths"] = r.evaluated_paths
    out["path_steps"] = r.path_steps
    out["time_steps"] = r.time_steps
    out["sobol_dimension"] = r.sobol_dimension
    out["device_ms"] = r.device_ms
    out["component_sum"] = list(r.component_sum)
    return out
