/*
 * pe.h -- C ABI of the MI355X Monte Carlo / RQMC pricing engine ("pe").
 *
 * The ABI replaces the simulation hot path behind the reference's facade:
 *
 *   PricingPlan::compile(&PricingRequest, ExecutionPolicy)
 *       crates/pricing/src/engine/risk/pricing.rs:53-88
 *   PricingPlan::evaluate(&self) -> MonteCarloPrice
 *       crates/pricing/src/engine/risk/pricing.rs:86-88, 395-404
 *   SimulationPlan::execute (the device boundary)
 *       crates/pricing/src/engine/risk/valuation.rs:19-35
 *
 * Everything crossing the boundary is plain data: POD structs, pointers and
 * sizes.  No exceptions cross it; every entry point returns a pe_status and
 * pe_last_error() returns a thread-local message for the last failure.
 * Dates are whole days counted from the valuation date (ACT/365F year fraction
 * = days / 365, crates/pricing/src/core/date.rs:223-231).
 *
 * Threading: like the reference's PricingPlan::evaluate (&self, Send + Sync),
 * distinct plans may be evaluated concurrently from several host threads.
 * Each evaluate call runs on a HIP stream of its own and holds the locks of
 * the plans it uses, so concurrent calls on the same plan serialise.  A plan
 * must not be freed while a call uses it.
 *
 * The oracle library (oracle/liboracle.so) accepts the same request structs
 * through oracle_evaluate(); it is test infrastructure, not part of this ABI.
 */
#ifndef PE_H
#define PE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PE_ABI_VERSION 3

/* Error taxonomy mirrors MonteCarloError (crates/pricing/src/api/error.rs:256-300)
 * at the granularity a C caller can act on. */
typedef enum pe_status {
    PE_OK = 0,
    PE_ERR_INVALID_ARGUMENT = 1,   /* request validation failed at compile */
    PE_ERR_UNSUPPORTED = 2,        /* product/model/engine combination not implemented */
    PE_ERR_NONFINITE_STATE = 3,    /* a sampling unit produced a non-finite state */
    PE_ERR_DEVICE = 4,             /* HIP runtime failure */
    PE_ERR_ALLOCATION = 5,         /* host or device allocation failed */
    PE_ERR_INSUFFICIENT_UNITS = 6, /* fewer than two independent units for a variance */
    PE_ERR_SIZE_LIMIT = 7          /* request exceeds a documented engine limit */
} pe_status;

typedef enum pe_backend {
    PE_BACKEND_GPU_FP64 = 1,  /* all-FP64 device mode: parity mode vs the oracle */
    PE_BACKEND_GPU_MIXED = 2  /* FP32 path state, FP64 accumulators */
} pe_backend;

/* ---------------------------------------------------------------- market -- */

/* LogLinearDiscountCurve (crates/pricing/src/market/curve.rs:39-212). */
typedef struct pe_curve {
    uint32_t count;
    const double* times;
    const double* discount_factors;
} pe_curve;

/* DividendEvent with DividendQuote::FixedCashAndProportional
 * (crates/pricing/src/market/dividend.rs:8-95; wire form DividendEventV1,
 * crates/pricing/src/wire.rs:480-492).  ex_time is an ACT/365F year fraction;
 * events are strictly ascending; fixed_cash >= 0, 0 <= beta < 1. */
typedef struct pe_dividend_quote {
    double ex_time;
    double fixed_cash;
    double beta;
} pe_dividend_quote;

typedef struct pe_market {
    double spot;
    pe_curve discount_curve;
    pe_curve dividend_curve;
    uint32_t dividend_count;
    const pe_dividend_quote* dividends;
} pe_market;

/* ----------------------------------------------------------------- model -- */

typedef enum pe_model_kind {
    PE_MODEL_BLACK_SCHOLES = 0,
    PE_MODEL_BLACK76 = 1,
    PE_MODEL_LOCAL_VOL = 2
} pe_model_kind;

/* LocalVarianceGrid (crates/pricing/src/market/local_variance.rs:128-135):
 * values are row-major [time][log-forward-moneyness]. */
typedef struct pe_local_variance_grid {
    uint32_t time_count;
    uint32_t x_count;
    const double* time_nodes;
    const double* x_nodes;
    const double* values;
    double floor;
    double cap;
} pe_local_variance_grid;

typedef struct pe_model {
    pe_model_kind kind;
    double volatility;              /* BS / Black-76 */
    pe_local_variance_grid lv;      /* local vol */
} pe_model;

/* --------------------------------------------------------------- product -- */

typedef enum pe_product_kind {
    PE_PRODUCT_EUROPEAN = 0,
    PE_PRODUCT_DIGITAL = 1,
    PE_PRODUCT_BARRIER = 2,
    PE_PRODUCT_ASIAN = 3,
    PE_PRODUCT_LOOKBACK = 4,
    PE_PRODUCT_AMERICAN = 5,    /* AmericanVanillaSpec: exercise dates in dates[], the last at expiry */
    PE_PRODUCT_GRAPH = 6        /* GraphProduct (crates/pricing/src/product/graph_product.rs:40-191):
                                   SSA payoff graph + cash flows in graph_nodes[] / graph_cashflows[];
                                   Black-Scholes / Black-76 models with pathwise Delta / Vega /
                                   Gamma (at most 20 live values with Greeks, else
                                   PE_ERR_UNSUPPORTED); local volatility price only (Greeks:
                                   PE_ERR_UNSUPPORTED); at most 32 values live at once after
                                   scheduling by observation (PE_ERR_SIZE_LIMIT); a live
                                   non-finite value or Greek fails the valuation
                                   (PE_ERR_NONFINITE_STATE) */
} pe_product_kind;

/* SourceOpcode (crates/pricing/src/product/graph.rs:89-215); the JSON names of
 * schemas/product-graph-v2/product_graph.schema.json in comments. */
typedef enum pe_graph_opcode {
    PE_GRAPH_LITERAL = 0,            /* literal: value */
    PE_GRAPH_TERMINAL_SPOT = 1,      /* terminal_spot: underlying_id, observation_day */
    PE_GRAPH_PRE_DIVIDEND_SPOT = 2,  /* pre_dividend_spot: underlying_id, observation_day */
    PE_GRAPH_ADD = 3,                /* add: left, right */
    PE_GRAPH_SUBTRACT = 4,           /* subtract: left, right */
    PE_GRAPH_MULTIPLY = 5,           /* multiply: left, right */
    PE_GRAPH_DIVIDE = 6,             /* divide: numerator, denominator */
    PE_GRAPH_MINIMUM = 7,            /* minimum: left, right */
    PE_GRAPH_MAXIMUM = 8,            /* maximum: left, right */
    PE_GRAPH_INDICATOR = 9,          /* indicator: input (1 if input >= 0) */
    PE_GRAPH_NEGATE = 10,            /* negate: input */
    PE_GRAPH_BARRIER_HIT_STATE = 11, /* barrier_hit_state: previous, hit_weight */
    PE_GRAPH_SMOOTH_MINIMUM = 12,    /* smooth_minimum: left, right, value = half width */
    PE_GRAPH_SMOOTH_MAXIMUM = 13,    /* smooth_maximum: left, right, value = half width */
    PE_GRAPH_SMOOTH_INDICATOR = 14,  /* smooth_indicator: input, value = half width */
    PE_GRAPH_LOG = 15,               /* log: input */
    PE_GRAPH_EXP = 16,               /* exp: input */
    PE_GRAPH_SELECT = 17,            /* select: condition, then_value, else_value */
    PE_GRAPH_COMPARE = 18            /* compare: left, right, comparison */
} pe_graph_opcode;

/* GraphComparison (graph.rs:68-87). */
typedef enum pe_graph_comparison {
    PE_GRAPH_LESS = 0,
    PE_GRAPH_LESS_EQUAL = 1,
    PE_GRAPH_EQUAL = 2,
    PE_GRAPH_NOT_EQUAL = 3,
    PE_GRAPH_GREATER_EQUAL = 4,
    PE_GRAPH_GREATER = 5
} pe_graph_comparison;

/* One source node (SourceNode, graph.rs:217-236).  operands[] hold node ids
 * in the order of the opcode's comment above; unused entries are ignored. */
typedef struct pe_graph_node {
    uint32_t id;
    pe_graph_opcode opcode;
    uint32_t operands[3];
    pe_graph_comparison comparison;
    uint32_t underlying_id;
    int32_t observation_day;         /* days after valuation */
    double value;                    /* literal value or smoothing half width */
} pe_graph_node;

/* Cash flow of a graph product: output node and its payment day. */
typedef struct pe_graph_cashflow {
    uint32_t output_node;
    int32_t payment_day;
} pe_graph_cashflow;

typedef enum pe_side { PE_CALL = 0, PE_PUT = 1 } pe_side;

/* ProductSpec (crates/pricing/src/product/spec.rs:14-140, 838-913).
 * dates[]: Asian observation dates, Barrier monitoring dates or Lookback
 * monitoring dates, ascending.  weights[]/fixings[] are per date for Asians
 * (fixings[i] is NaN for an unknown observation). */
typedef struct pe_product {
    pe_product_kind kind;
    pe_side side;
    int32_t expiry_day;
    int32_t payment_day;
    double strike;
    double notional;
    /* Digital */
    double payout;
    int32_t payout_is_asset;
    /* Barrier */
    double barrier;
    int32_t direction_down;     /* 0 = up, 1 = down */
    int32_t knock_in;           /* 0 = knock-out, 1 = knock-in */
    int32_t continuous;         /* 0 = discrete monitoring, 1 = continuous */
    int32_t has_rebate;
    double rebate;
    int32_t historical_hit;     /* -1 = none, 0 = not hit, 1 = hit */
    /* Lookback */
    int32_t has_historical_extremum;
    double historical_extremum;
    /* schedules */
    uint32_t date_count;
    const int32_t* dates;
    const double* weights;
    const double* fixings;
    /* Graph (PE_PRODUCT_GRAPH): nodes, cash flows and the declared
     * underlyings; observations must be on the first (the market's). */
    uint32_t graph_node_count;
    const pe_graph_node* graph_nodes;
    uint32_t graph_cashflow_count;
    const pe_graph_cashflow* graph_cashflows;
    uint32_t graph_underlying_count;
    const uint32_t* graph_underlyings;
} pe_product;

/* ---------------------------------------------------------------- engine -- */

typedef enum pe_engine_kind {
    PE_ENGINE_PSEUDO = 0, /* PseudoMcConfig (crates/pricing/src/engine/mc/config.rs) */
    PE_ENGINE_RQMC = 1    /* RqmcConfig */
} pe_engine_kind;

typedef struct pe_engine {
    pe_engine_kind kind;
    uint64_t seed;              /* master seed / master scramble seed */
    uint64_t units;             /* pseudo: independent units; RQMC: points per scramble */
    uint32_t scrambles;         /* RQMC only */
    int32_t antithetic;
    int32_t brownian_bridge;
} pe_engine;

/* RiskRequest (crates/pricing/src/risk/request.rs). */
typedef struct pe_risk {
    int32_t delta;
    int32_t vega;
    int32_t gamma;              /* 0 = off */
    int32_t gamma_relative;     /* 1: bump = spot * gamma_bump, 0: absolute */
    double gamma_bump;
    int32_t smoothing;          /* 1: PayoffSmoothing::CompactC2 */
    double smoothing_half_width;
    /* RiskRequest aad_tile_capacity / checkpoint_interval overrides (0: the
     * reference's defaults, min(256, block size) and 64).  The device has no
     * AAD tape; they are validated and echoed in pe_mc_diagnostics. */
    uint32_t aad_tile_capacity;
    uint32_t checkpoint_interval;
} pe_risk;

/* Continuation-value regression of an early-exercise product. */
typedef enum pe_lsm_method {
    PE_LSM_POLY_CPQR = 0,           /* the reference's column-pivoted Householder QR on
                                       standardized monomials (crates/pricing/src/engine/mc/lsm.rs:1928-2568) */
    PE_LSM_POLY_LAGUERRE = 1,       /* weighted Laguerre basis, Gram + FP64 Cholesky (no reference counterpart) */
    PE_LSM_MLP = 2                  /* MLP continuation network, BF16 weights (no reference counterpart) */
} pe_lsm_method;

/* LsmConfig (crates/pricing/src/engine/mc/lsm.rs, api request "lsm"): the
 * training engine (domain LsmTrain for pseudo-MC), the spot-feature basis
 * PolynomialBasisSpec(1, max_degree), CPQR rank tolerances, the ITM
 * tolerance and the regression matrix element limit; plus the device
 * method and the MLP hyper-parameters. */
typedef struct pe_lsm {
    pe_engine training_engine;
    uint32_t max_degree;
    double abs_rank_tolerance;
    double rel_rank_tolerance;
    double itm_abs_tolerance;
    uint64_t max_matrix_elements;
    pe_lsm_method method;
    uint32_t mlp_hidden;            /* hidden width (64) */
    uint32_t mlp_epochs;            /* Adam epochs per exercise date (4) */
    uint32_t mlp_batch;             /* mini-batch rows (8192) */
    double mlp_learning_rate;       /* Adam step (1e-3) */
    uint64_t mlp_seed;              /* initialisation and shuffling */
} pe_lsm;

typedef struct pe_request {
    pe_market market;
    pe_model model;
    pe_product product;
    pe_engine engine;
    pe_risk risk;
    pe_lsm lsm;                     /* early-exercise products only */
} pe_request;

/* ExecutionPolicy (crates/pricing/src/engine/mc/execution.rs:20-68). */
typedef struct pe_exec_policy {
    uint32_t worker_threads;
    uint64_t reduction_block_size;   /* 0 selects the default 4096 */
} pe_exec_policy;

/* ---------------------------------------------------------------- result -- */

/* Estimate (crates/pricing/src/engine/risk/report.rs:454-473). */
typedef struct pe_estimate {
    double value;
    double standard_error;
    double confidence_lower;
    double confidence_upper;
    uint64_t effective_units;
} pe_estimate;

/* Component order of the 18-wide pathwise vector
 * (crates/pricing/src/engine/plan/simulation.rs:26-62). */
enum {
    PE_C_PRICE = 0, PE_C_DELTA = 1, PE_C_VEGA = 2, PE_C_GAMMA = 3,
    PE_C_BUMP_DELTA = 4, PE_C_DELTA_DIFFERENCE = 5, PE_C_BUMP_VEGA = 6,
    PE_C_VEGA_DIFFERENCE = 7, PE_C_BUMP_GAMMA = 8, PE_C_GAMMA_DIFFERENCE = 9,
    PE_C_BARRIER_ENDPOINT_HIT = 10, PE_C_BARRIER_DIVIDEND_JUMP_HIT = 11,
    PE_C_BARRIER_BRIDGE_HIT_WEIGHT = 12, PE_C_BARRIER_INTERVAL_COUNT = 13,
    PE_C_BARRIER_FINITE_CORRECTION_COUNT = 14, PE_C_BARRIER_ZERO_VARIANCE_COUNT = 15,
    PE_C_BARRIER_SURVIVAL_UNDERFLOW_COUNT = 16, PE_C_BARRIER_CERTAIN_SURVIVAL_COUNT = 17,
    PE_COMPONENTS = 18
};

/* EstimatorKind, CurveRegion (crates/pricing/src/market/curve.rs:18-22) and
 * PayoffValuationKind (crates/pricing/src/api/mc_result.rs:75-78). */
typedef enum pe_estimator { PE_ESTIMATOR_PSEUDO_MONTE_CARLO = 0, PE_ESTIMATOR_RQMC = 1 } pe_estimator;
typedef enum pe_curve_region {
    PE_CURVE_PILLAR = 0, PE_CURVE_INTERPOLATED = 1, PE_CURVE_RIGHT_EXTRAPOLATED = 2
} pe_curve_region;
typedef enum pe_valuation_kind {
    PE_VALUATION_EXACT_CONTRACTUAL = 0, PE_VALUATION_SMOOTHED_SURROGATE = 1
} pe_valuation_kind;

/* MonteCarloDiagnostics (crates/pricing/src/api/mc_result.rs:226-247) except
 * the payoff-graph fingerprint and the optional smoothing / path-state /
 * bridge sub-blocks (the bridge diagnostics are result components 10-17).
 * RQMC plans carry the BLAKE3-256 direction checksum (of the Joe-Kuo asset)
 * and scramble checksum (of every scramble's digital shift and matrix rows,
 * big-endian, scramble-major; crates/pricing/src/engine/sampling/qmc.rs:236-259).
 * policy_version, aad_tile_policy_version and checkpoint_policy_version are
 * the reference's policy versions (1). */
typedef struct pe_mc_diagnostics {
    uint64_t master_seed;
    pe_estimator estimator;
    uint32_t scramble_count;            /* RQMC; 0 = none (pseudo-MC) */
    int32_t has_checksums;              /* 1 for RQMC */
    uint8_t direction_checksum[32];
    uint8_t scramble_checksum[32];
    uint32_t policy_version;
    uint32_t worker_threads;
    uint64_t reduction_block_size;
    uint32_t aad_tile_policy_version;
    uint32_t aad_tile_capacity;
    uint32_t checkpoint_policy_version;
    uint32_t checkpoint_interval;
    int32_t antithetic;
    pe_curve_region discount_region;    /* discount curve at the payment time */
    pe_curve_region dividend_region;    /* dividend curve at the expiry (0 without observations) */
    pe_valuation_kind valuation_kind;
} pe_mc_diagnostics;

typedef struct pe_result {
    /* per component, over independent units (pseudo) or scramble means (RQMC) */
    pe_estimate estimates[PE_COMPONENTS];
    double component_sum[PE_COMPONENTS];      /* Neumaier total */
    double component_mean[PE_COMPONENTS];     /* centred-moment mean */
    double component_m2[PE_COMPONENTS];       /* centred second moment */
    double sampling_variance;
    double estimator_variance;
    uint64_t independent_units;
    uint64_t evaluated_paths;
    uint64_t path_steps;                      /* simulated (path, node) updates */
    uint32_t time_steps;                      /* steps per path */
    uint32_t sobol_dimension;                 /* RQMC effective dimension */
    double device_ms;                         /* kernel time, HIP events */
    uint64_t failing_unit;                    /* first failing sampling unit */
    pe_mc_diagnostics diagnostics;            /* of the valuation engine */
} pe_result;

/* ------------------------------------------------------------ entry points -- */

typedef struct pe_plan pe_plan;

/* Library bring-up; idempotent.  Selects HIP device `device`. */
pe_status pe_init(int32_t device);
const char* pe_last_error(void);
int32_t pe_abi_version(void);

/* Mirrors SimulationPlan::compile (crates/pricing/src/engine/compile/simulation.rs:78-640):
 * host FP64 plan compilation, flattening to device tables, upload. */
pe_status pe_plan_compile(const pe_request* request, const pe_exec_policy* policy,
                          pe_plan** out);
void pe_plan_free(pe_plan* plan);

/* The MonteCarloDiagnostics a plan's evaluation reports, known at compile
 * time (host only; every evaluate copies it into pe_result.diagnostics). */
pe_status pe_plan_diagnostics(const pe_plan* plan, pe_mc_diagnostics* out);

/* BLAKE3-256 of `bytes` bytes (the checksum primitive above; exposed for
 * verification, no reference counterpart). */
pe_status pe_blake3_256(const void* data, size_t bytes, uint8_t out[32]);

/* Mirrors PricingPlan::evaluate (crates/pricing/src/engine/risk/pricing.rs:86-88). */
pe_status pe_plan_evaluate(const pe_plan* plan, pe_backend backend, pe_result* out);

/* Allocation-free variant of pe_plan_evaluate for Black-Scholes / Black-76
 * plans (price, Greeks, continuous barriers, graph products; RQMC or
 * pseudo-MC).  No reference counterpart: it is for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Every device buffer of the
 * evaluation (trade record, observation and Sobol tables, graph tape, work
 * items, block partials, error word) lives in one fixed 8 MiB workspace per
 * GPU, a file-scope device array held until the process ends; a plan that
 * needs more returns PE_ERR_SIZE_LIMIT (roughly: 4096-unit chunks x
 * scrambles x 40 B per statistic, plus the Sobol table of 132 B per
 * dimension and scramble).  The call makes no hipMalloc/hipFree, pinned-host
 * allocation or event creation; host-side bookkeeping uses the C++ heap.
 * All calls on one GPU share its workspace: they are serialised inside the
 * library, and the workspace keeps the last call's contents.  Local-vol,
 * lifted Heston and American plans return PE_ERR_UNSUPPORTED.  device_ms
 * is 0 (no timing events).  Results equal pe_plan_evaluate's. */
pe_status pe_plan_evaluate_noalloc(const pe_plan* plan, pe_backend backend, pe_result* out);

/* Evaluates several plans that share one simulated path set (same market,
 * model, engine and simulation grid; e.g. a barrier and an Asian on the same
 * local-vol paths) in one launch.  Each result equals what pe_plan_evaluate
 * returns for that plan.  No reference counterpart: the reference compiles one
 * plan per request and simulates each separately. */
pe_status pe_plans_evaluate_shared(const pe_plan* const* plans, uint32_t count,
                                   pe_backend backend, pe_result* out);

/* Evaluates `count` independent Black-Scholes / Black-76 plans (thousands of
 * trades with uneven step and path counts) in persistent launches that share
 * the device: results[i] equals what pe_plan_evaluate returns for plans[i].
 * Each plan follows SimulationPlan::execute
 * (crates/pricing/src/engine/risk/valuation.rs:19-35); the batch has no
 * reference counterpart beyond calling price_monte_carlo once per trade
 * (crates/pricing/src/engine/risk/valuation.rs:440-455).  A path that fails
 * fails the call; the message names the trade and the sampling unit. */
pe_status pe_batch_evaluate(const pe_plan* const* plans, uint32_t count, pe_backend backend,
                            pe_result* results);

/* --------------------------------------------------------- early exercise -- */

/* One exercise date's fitted continuation model (ExerciseDecisionModel,
 * crates/pricing/src/engine/mc/lsm.rs:776-819 and its regression diagnostics,
 * :2067-2220).  coefficients are per basis column in the original order
 * (exponent k of the standardized spot, or weighted Laguerre function k);
 * diagonal_abs are |R_pp| in pivot order. */
typedef struct pe_lsm_decision {
    int32_t continue_all;           /* no in-the-money training row */
    uint32_t rank;
    uint64_t itm_rows;
    double coefficients[6];         /* per basis column (spot policies: 4; spot and variance: 1 s v s^2 sv v^2) */
    double diagonal_abs[6];
    uint32_t pivot_order[6];        /* original column per pivot position (active columns) */
    double feature_mean;            /* FeatureScaling of the spot (monomial basis) */
    double feature_population_variance;
    double feature_scale;
    int32_t feature_inactive;
    uint32_t active_columns;        /* pivot entries in use */
    double rank_threshold;
    double residual_sum_squares;
    /* MLP policy: the exercise region in moneyness S/K at this date, for
     * in-the-money paths: inside iff region_left xor (number of cuts <= S/K
     * is odd) */
    int32_t region_left;
    uint32_t region_cut_count;
    float region_cuts[64];
    double zero_scale_threshold;    /* FeatureScaling: 64 eps max(1, max |ITM spot|) (monomial basis) */
    /* lifted-Heston policies: FeatureScaling of the variance state */
    double variance_mean;
    double variance_population_variance;
    double variance_scale;
    double variance_zero_scale_threshold;
    int32_t variance_inactive;
    uint32_t feature_count;         /* 1: spot; 2: spot and variance */
} pe_lsm_decision;

/* Early-exercise diagnostics of one plan (EarlyExerciseDiagnostics,
 * crates/pricing/src/engine/risk/early_exercise.rs:24-300): the in-sample
 * value of the trained policy, valuation exercise counts per date and the
 * decision model per non-terminal date; plus device time per phase. */
typedef enum pe_random_domain {     /* RandomDomain (crates/pricing/src/engine/sampling/random.rs:13-20) */
    PE_DOMAIN_VALUATION = 0, PE_DOMAIN_LSM_TRAIN = 1, PE_DOMAIN_RQMC_SCRAMBLE = 2
} pe_random_domain;

typedef struct pe_lsm_report {
    double in_sample_value;
    double training_ms;             /* training paths + backward induction, HIP events */
    double valuation_ms;            /* out-of-sample valuation */
    uint32_t exercise_dates;
    uint32_t capacity;              /* entries of the caller arrays below */
    uint64_t* exercise_counts;      /* [capacity] or NULL */
    pe_lsm_decision* decisions;     /* [capacity] or NULL */
    double path_ms;                 /* training-path generation, part of training_ms */
    /* EarlyExerciseDiagnostics (crates/pricing/src/api/mc_result.rs:249-273),
     * all but the policy fingerprint and the regression warnings. */
    double* exercise_probabilities; /* [capacity] or NULL: counts / valuation trajectories */
    int32_t* dividend_collisions;   /* [capacity] or NULL: an ex-dividend time equals the date's */
    uint32_t* stopping_indices;     /* [stopping_capacity] or NULL: exercise-date index per valuation
                                       trajectory, (scramble, unit, antithetic pass) order */
    uint64_t stopping_capacity;
    pe_random_domain training_random_domain;
    pe_random_domain valuation_random_domain;
    uint64_t training_sampling_units, training_trajectories;
    uint64_t valuation_sampling_units, valuation_trajectories;
    int32_t training_has_checksums, valuation_has_checksums;
    uint8_t training_direction_checksum[32], training_scramble_checksum[32];
    uint8_t valuation_direction_checksum[32], valuation_scramble_checksum[32];
    uint32_t policy_feature_count, policy_max_degree;   /* PolynomialBasisSpec */
    double itm_abs_tolerance, abs_rank_tolerance, rel_rank_tolerance;
    uint64_t max_matrix_elements;
    /* Network policy (PE_LSM_MLP) only; no reference counterpart.  The
     * weighted-Laguerre polynomial policy is trained on the same training
     * paths and valued on the same valuation paths as the network: it is the
     * judge, and the fallback when its out-of-sample value is the higher one
     * (fallback_used = 1: the result, exercise counts and probabilities are
     * the polynomial policy's).  Stopping indices and the decisions' region
     * cuts are the network's; the decisions' coefficients are the judge's. */
    pe_estimate network_price;      /* out-of-sample value of the network policy */
    pe_estimate judge_price;        /* out-of-sample value of the polynomial policy */
    double judge_in_sample_value;
    int32_t fallback_used;
    uint32_t pad;
    double network_training_ms;     /* training paths + network induction (HIP events) */
    double judge_training_ms;       /* the polynomial judge's induction on the same paths */
    double network_valuation_ms, judge_valuation_ms;
    uint64_t network_rows_trained;  /* rows run through the network: epochs x in-the-money rows, all dates */
} pe_lsm_report;

/* Prices `count` American plans by least-squares Monte Carlo
 * (execute_early_exercise / execute_early_exercise_rqmc,
 * crates/pricing/src/engine/risk/early_exercise.rs:24-300): training paths
 * in the LsmTrain domain, backward induction, and an out-of-sample valuation
 * on independent paths, all on the device.  results[i] equals what
 * pe_plan_evaluate returns for plans[i]; reports (optional, `count` entries)
 * receive the policy diagnostics.  The method of each plan's pe_lsm selects
 * the reference's monomial CPQR policy (solved by normal equations with its
 * pivot and rank rules), the weighted-Laguerre policy or the MLP policy (one
 * network per trade, judged against and falling back to the weighted-Laguerre
 * policy; price-only: fixed-policy Greeks are PE_ERR_UNSUPPORTED). */
pe_status pe_lsm_evaluate(const pe_plan* const* plans, uint32_t count, pe_backend backend, pe_result* results,
                          pe_lsm_report* reports);

/* Andersen-Broadie primal-dual bounds of one American plan
 * (AndersenBroadiePlan::compile / evaluate,
 * crates/pricing/src/engine/risk/dual.rs:143-296): the plan's LSM policy is
 * trained as for pricing, the plan's valuation engine supplies the outer
 * paths, and from every outer path and non-terminal exercise date the policy
 * drives `continuation_inner_paths` (continuation region) or
 * `exercise_inner_paths` (exercise region) independent inner rollouts in a
 * private random domain seeded by `inner_seed`.  Price-only Black-Scholes /
 * Black-76 plans with pseudo-MC training and valuation engines and a
 * polynomial policy; FP64 backend. */
typedef struct pe_dual_result {
    pe_estimate lower_bound;
    pe_estimate upper_bound;
    pe_estimate duality_gap;
    double price_confidence_interval_95[2];   /* [lower - 1.96 SE, upper + 1.96 SE] */
    uint64_t outer_trajectories;
    uint32_t exercise_date_count;
    uint32_t pad;
    double device_ms;
} pe_dual_result;

pe_status pe_lsm_dual_evaluate(const pe_plan* plan, uint32_t continuation_inner_paths,
                               uint32_t exercise_inner_paths, uint64_t inner_seed, pe_backend backend,
                               pe_dual_result* out);

/* Continuation network of the MLP policy, exposed for verification (no
 * reference counterpart).  Parameters: 4417 floats, W1[64][2] | b1[64] |
 * W2[64][64] (row = output unit) | b2[64] | w3[64] | b3; inputs x row-major
 * rows x 2.  pe_mlp_gradients returns the gradient of the mean squared error
 * (1/rows) sum (out - y)^2 and the loss, computed by the device's BF16
 * matrix-core forward and backward passes with FP32 accumulation. */
pe_status pe_mlp_gradients(const float* params, const float* x, const float* y, uint32_t rows, float* grad_out,
                           double* loss_out);
pe_status pe_mlp_forward(const float* params, const float* x, uint32_t rows, float* out);

/* ------------------------------------------------- lifted Heston (rough) -- */

/* LiftedHeston (crates/pricing/src/models/rough_volatility.rs:172-212): a
 * finite exponential-kernel Heston, V = v0 + sum_i w_i f_i, every factor
 * driven by one variance Brownian motion; rates x_i in inverse years. */
typedef struct pe_lifted_heston {
    double initial_variance;
    double mean_reversion;
    double long_run_variance;
    double vol_of_vol;
    double correlation;
    uint32_t factor_count;          /* 1..256 */
    const double* weights;
    const double* rates;
} pe_lifted_heston;

/* LiftedHeston::from_rough (crates/pricing/src/models/rough_volatility.rs:564-606):
 * Abi Jaber's geometric-bin weights and rates of the fractional kernel for
 * Hurst index `hurst`, `factors` factors and bin ratio `ratio` (> 1). */
pe_status pe_lifted_heston_from_rough(double hurst, uint32_t factors, double ratio, double* weights_out,
                                      double* rates_out);

/* RoughVolatilityPricingPlan::compile (crates/pricing/src/engine/risk/rough_volatility.rs:84-190)
 * for a lifted Heston model.  The request is a Black-Scholes carrier of market,
 * payoff, dates and engine (its volatility is not used); risk must be off and
 * the market must have no discrete dividends.  The time grid is the
 * local-volatility grid of the events {0, expiry, fixings} with at most
 * `maximum_step` between nodes.  Evaluate with pe_plan_evaluate (the
 * reference's per-plan grid) or pe_plans_evaluate_shared (several plans on one
 * simulated path set over the union of their events; no reference
 * counterpart). */
pe_status pe_rough_plan_compile(const pe_request* carrier, const pe_lifted_heston* model, double maximum_step,
                                const pe_exec_policy* policy, pe_plan** out);
/* An American carrier (PE_PRODUCT_AMERICAN: a Bermudan on its exercise
 * dates) under lifted Heston is priced by least squares with the features
 * (spot, variance state): the reference's regression (standardized
 * monomials of total degree lsm.max_degree <= 2, CPQR rules) on in-the-money
 * rows, exercising in-the-money paths only; training paths in FP64 in both
 * modes.  Evaluate with pe_plan_evaluate or pe_lsm_evaluate (one plan per
 * launch).  No reference counterpart: the reference rejects early exercise
 * in its hybrid adapters (crates/pricing/src/engine/compile/simulation.rs:55-59). */

/* LocalVarianceGrid::from_surface for a StandardSsvi surface with a ThetaPchip
 * total-variance curve and a power-law phi, computed in FP64 on the device
 * (crates/pricing/src/market/local_variance.rs:149-206, ssvi.rs:61-186,226-275,
 * 340-475, pchip.rs:111-170).  values_out is row-major [time][x];
 * repair_count_out receives the number of floor/cap repairs. */
typedef struct pe_ssvi_surface {
    uint32_t theta_count;
    const double* theta_times;
    const double* theta_values;
    double theta_terminal_slope;
    double rho;
    double eta;
    double gamma;
} pe_ssvi_surface;

pe_status pe_local_variance_from_ssvi(const pe_ssvi_surface* surface,
                                      uint32_t time_count, const double* time_nodes,
                                      uint32_t x_count, const double* x_nodes,
                                      double floor, double cap,
                                      double* values_out, uint32_t* repair_count_out);

/* Accuracy of the mixed mode's FP32 inverse normal on the device: the largest
 * relative error of z against the FP64 AS241 over words [first_word,
 * first_word + count) and a word attaining it (verification, no reference
 * counterpart). */
pe_status pe_normal_mixed_max_error(uint64_t first_word, uint64_t count, double* max_relative_error,
                                    uint32_t* worst_word);

/* Kernel-time instrumentation of the last evaluate on this thread: average
 * duration of the dominant kernel over its launches, in milliseconds. */
pe_status pe_last_kernel_stats(double* avg_ms, uint32_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* PE_H */
