// Plain-data descriptors shared by the host plan compiler and the kernels.
// Everything here is POD so it can be passed by value as a kernel argument or
// copied to HBM as a table.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace pe {

constexpr int kMaxProducts = 4;      // products sharing one simulated path set
constexpr int kMaxVariants = 5;      // base, validation (= Gamma) spot +/-, vol +/-
constexpr int kMaxBridgeDepth = 64;  // streaming Brownian-bridge stack (LDS)
constexpr int kBlockThreads = 256;
constexpr int kComponents = 18;      // the 18-wide pathwise risk vector (PE_COMPONENTS)

enum ProductKind : int32_t { kEuropean = 0, kDigital = 1, kBarrier = 2, kAsian = 3, kLookback = 4, kGraph = 5 };

// ----------------------------------------------------------- graph payoffs --
// A graph product's compiled tape (CompiledPayoff, crates/pricing/src/engine/
// payoff/tape.rs:8-143) as a register program: operations in observation
// order (an operation runs right after the latest observation it depends
// on), values in kGraphMaxRegs registers per lane, constants in a table.
// An operand below kGraphConst is a register, otherwise a constant index.
constexpr uint32_t kGraphMaxRegs = 32;
// with pathwise Greeks each register also carries two tangents in LDS
constexpr uint32_t kGraphMaxRiskRegs = 20;
constexpr uint16_t kGraphConst = 0x8000u;
// operands read straight from the current observation (a spot node whose
// every consumer runs in its own observation's segment is not materialised)
constexpr uint16_t kGraphPost = 0x4000u;
constexpr uint16_t kGraphPre = 0x4001u;

enum GraphCode : uint16_t {
    kGSpot = 0, kGPreSpot, kGAdd, kGSub, kGMul, kGDiv, kGMin, kGMax, kGInd, kGNeg, kGHit,
    kGSmoothMin, kGSmoothMax, kGSmoothInd, kGLog, kGExp, kGSelect,
    kGLess, kGLessEqual, kGEqual, kGNotEqual, kGGreaterEqual, kGGreater,
    kGMulAdd   // a * b + c: a Multiply fused into the Add that is its only consumer
};

struct alignas(64) GraphOp {   // one 64-byte scalar load per operation
    uint32_t code, out;      // GraphCode, destination register
    uint32_t a, b, c;        // operands (register, or kGraphConst: value in ka/kb/kc)
    uint32_t pad;
    double param;            // smoothing half width
    double ka, kb, kc;       // constant operand values
    double pad2;
};

// One product's payoff, specialised per family (the reference builds these as
// SSA graphs, crates/pricing/src/engine/payoff/compile.rs:35-474).
struct ProductDesc {
    int32_t kind;
    int32_t call;            // 1 call, 0 put
    int32_t payout_asset;    // digital asset-or-nothing
    int32_t down;            // barrier direction
    int32_t knock_in;
    int32_t has_rebate;
    int32_t smoothing;       // compact-C2 smoothing of indicators
    int32_t initial_hit;     // barrier: historical hit -> 1
    int32_t has_hist_ext;    // lookback historical extremum
    int32_t terminal_obs;    // observation index of the expiry fixing
    int32_t first_obs;       // lookback: first future observation
    int32_t continuous;      // barrier monitored continuously (bridge survival)
    double strike, notional, payout, barrier, rebate, half_width;
    double discount;         // DF(payment)
    double hist_ext;         // lookback historical extremum
    double asian_prefix;     // Asian: known-fixing prefix sum
    float f_barrier, f_half_width, f_strike, f_pad;  // FP32 copies for the mixed kernels
};

// One observation of a BS/Black-76 trade: time, Brownian step, forward and
// the escrowed-dividend affine coordinates S = A*S0 + B*F after (and, on an
// ex-date, before) the dividend (crates/pricing/src/market/dividend.rs:136-205,
// crates/pricing/src/engine/processes/black_scholes.rs:335-354).
enum BsObsFlags : uint32_t {
    kObsPre = 1u,       // pre-dividend coordinate present (ex-date)
    kObsFixing = 2u,    // fixing of the product (Asian weight, barrier or lookback date)
    kObsJump = 4u,      // discrete barrier monitors the pre-dividend spot too
    kObsBridge = 8u,    // continuous barrier: bridge observation
};

struct BsObs {
    double time, sqrt_step, forward, weight;
    double a, b, ss;          // post-dividend coordinate (A, B, spot scale)
    double pa, pb, pss;       // pre-dividend coordinate
    // the coordinates refitted at the spot-bumped risk variants 1 and 2
    // (AffineDividendCoordinate::at_spot; the other variants use the base
    // spot): post (va, vb) and pre (vpa, vpb), spot scales unchanged
    double va[2], vb[2], vpa[2], vpb[2];
    uint32_t flags;
    uint32_t pad;
};

// A pathwise evaluation point (spot, vol) of the BS risk vector
// (crates/pricing/src/engine/processes/black_scholes.rs:417-472).
struct Variant {
    double spot;
    double vol;
    double spot_scale;   // spot / S0
    float f_spot, f_vol;  // FP32 copies for the mixed kernels
    float f_vol2;         // vol log2(e): the exponent's Brownian term in base 2
    float f_inv_spot;     // 1 / spot
};

// FP32 per-observation constants of the mixed batch kernel, computed in FP64
// and rounded once.  Spot variants si: 0 base, 1 S - h, 2 S + h (the risk
// variants 1 and 2); vol variants vi: 0 base, 1 sigma - hv, 2 sigma + hv
// (the risk variants 3 and 4); every other risk variant uses index 0.
struct BsObs32 {
    float time, sqrt_step, weight, sqrt_time;
    float fwd[3];            // forward x spot scale
    float cash[3], b[3];     // post-dividend S = cash + B F: cash = A spot
    float pcash[3], pb[3];   // pre-dividend
    float ssos[3], pssos[3]; // spot scale / spot (d S / d S0 per unit F), post and pre
    float ss, pss;
    float a[3], pa[3];       // A at the spot variants (continuous barrier states)
    float bar[3], pbar[3];   // (barrier - A spot) / B, post and pre (continuous barrier)
    float mh2[3];            // -sigma^2 t / 2 log2(e)
    float nvt[3];            // -sigma t
    float sd2[3];            // sigma sqrt(t) log2(e) (single terminal step)
    // continuous barrier, per vol variant: the variance of the bridge interval
    // ending here (sigma^2 times the time since the previous bridge
    // observation), its log and reciprocal; per spot variant: 1 / (B x the
    // transformed barrier), post and pre
    float iv[3], liv[3], iiv[3];
    float itsb[3], pitsb[3];
    uint32_t flags;          // BsObsFlags
    float pad;
};

// Streaming bridge program: PUSH computes the value at a split node from the
// current left value and the stack top; POP emits the next grid node.
struct BridgeOp {
    int32_t push;        // 1 push, 0 pop
    int32_t rank;        // normal rank (Sobol / Philox dimension)
    int32_t terminal;    // push of the terminal node
    int32_t step;        // pop: the step whose right node is emitted
    double left_weight, right_weight, sd;
    float f_left_weight, f_right_weight, f_sd, pad;
};

// Per-step constants of the LV Log-Euler scheme, all computed on the host in
// FP64 with the reference's expressions
// (crates/pricing/src/engine/processes/local_vol.rs:491-513).
struct LvStep {
    double t, dt, sqrt_dt, ratio, inv_sqrt_dt, forward, next_forward;
    int32_t t_cell;
    int32_t node_mask;   // products observing the step's right node
    double t_weight;
    // FP32 copies for the mixed kernel (no per-step conversions on the device)
    float m_half_dt;     // -dt / 2
    float f_sqrt_dt;
    float f_t_weight;
    float f_inv_sqrt_dt;
    float f_next_forward;
    float pad32[3];
};

// One instruction of the local-vol program: a bridge push, a bridge pop that
// emits step `index`, or (no bridge) a direct step whose shock is normal
// `index`.  Every constant the step needs travels with it, so one prefetched
// scalar load per instruction feeds the loop.  code packs
//   bits 0-1 kind, bit 2 terminal push, bits 4-7 product mask of the emitted
//   node, bits 8-31 time cell of the local-variance grid.
enum LvOpKind : uint32_t { kLvPush = 0, kLvPop = 1, kLvDirect = 2 };

__host__ __device__ inline uint32_t lv_code(uint32_t kind, uint32_t terminal, uint32_t mask, uint32_t t_cell) {
    return kind | (terminal << 2) | (mask << 4) | (t_cell << 8);
}

struct LvOp32 {          // mixed kernel (12 dwords)
    uint32_t code;
    int32_t index;
    float lw, rw, sd, inv_sqrt_dt;
    float m_half_dt, sqrt_dt, t_weight, next_forward;
    float pad0, pad1;
};
// An RQMC bridge program has one copy of the stream per scramble; in each
// copy a push's dwords 6-11 (step fields, unused by a push) hold that
// scramble's direction words 0-5 of the push's Sobol dimension, so the
// push's record brings the lane part of its Sobol word in the same load.
constexpr int kLvPushDirWord = 6;
static_assert(offsetof(LvOp32, m_half_dt) == kLvPushDirWord * 4 && offsetof(LvOp32, pad1) == (kLvPushDirWord + 5) * 4 &&
                  sizeof(LvOp32) == 48,
              "a push's direction words fill dwords 6-11");

struct LvOp64 {          // FP64 kernel (the reference's constants, bit for bit)
    uint32_t code;
    int32_t index;
    double lw, rw, sd, inv_sqrt_dt;
    double dt, sqrt_dt, ratio, forward, t_weight;
};

// Hot per-product state of the local-vol kernel: what a fixing does.
enum LvMode : int32_t {
    kModeTerminal = 0,     // European / digital: expiry spot only
    kModeAsian = 1,        // acc += S * w
    kModeBarrierUp = 2,    // hard indicator, hit is absorbing
    kModeBarrierDown = 3,
    kModeBarrierSmooth = 4,
    kModeLookbackMax = 5,
    kModeLookbackMin = 6,
};

struct ProductHot {
    int32_t mode;
    int32_t terminal_node;
    float barrier;
    float half_width;
};

// One BS/Black-76 trade of a batch (one compiled plan).
struct BsTrade {
    ProductDesc product;
    Variant variants[kMaxVariants];
    uint64_t seed;
    uint64_t units;            // pseudo units or RQMC points per scramble
    double s0;
    double reserve;            // escrowed cash reserve (0 without cash dividends)
    double init_a, init_b;     // initial affine coordinate (continuous barrier)
    double gamma_bump, spot_bump, vol_bump;
    double resolved_survival;  // continuous barrier decided by history
    uint32_t obs_offset, n_obs;
    uint32_t sobol_offset;     // words into the batch Sobol table
    uint32_t sobol_dims;
    uint32_t n_variants;
    uint32_t rqmc, antithetic, single_terminal;
    uint32_t request_gamma, zero_variance;
    uint32_t has_dividends;    // any non-identity coordinate
    uint32_t continuous;       // continuous barrier monitoring
    uint32_t resolved;         // continuous barrier: survival fixed by history
    uint32_t expiry_obs;
    uint32_t risk;
    // graph product: ops in [graph_ops + seg[0], ...); seg[i + 1] begins the
    // ops that run after observation i (n_obs + 2 boundaries from graph_segs)
    uint32_t graph_ops, graph_segs, graph_consts;
    uint16_t graph_output, graph_regs;
};

// A unit range of one trade and scramble, reduced by one workgroup into one
// BlockStat per component at `partial`.
struct BsItem {
    uint32_t trade;
    uint32_t scramble;
    uint64_t unit_begin, unit_end;
    uint32_t partial;
    uint32_t pad;
};

struct BsBatchArgs {
    const BsTrade* trades;
    const BsObs* obs;
    const BsObs32* obs32;      // the same observations' FP32 constants (mixed kernels)
    const uint32_t* sobol;
    const BsItem* items;
    void* partials;            // BlockStat [partial][component]
    unsigned long long* error; // smallest (trade << 40 | unit) whose path failed
    const GraphOp* graph_ops;  // graph products: tapes, boundaries, constants
    const uint32_t* graph_segs;
    const double* graph_consts;
    uint32_t n_items;
    uint32_t components;       // statistics per item (1, 10 or 18)
    uint32_t grid;             // persistent workgroups (0: one per item)
    uint32_t graph_regs;       // graph groups: LDS registers per lane
};

// ------------------------------------------------------- lifted Heston --
constexpr int kMaxRoughFactors = 32;

// Factor slots of the kernel instantiation that runs `factors` factors (8,
// 16, 20 or 32; 0 above kMaxRoughFactors): the [step][factor] tables are laid
// out with this stride, padded with factors that stay at zero.
__host__ __device__ constexpr uint32_t rough_factor_slots(uint32_t factors) {
    return factors <= 8 ? 8u : factors <= 16 ? 16u : factors <= 20 ? 20u : factors <= 32 ? 32u : 0u;
}

struct RoughStep {       // per time step of the shared grid
    double dt, sqrt_dt;
    double scale_next;   // physical spot per unit of F at the step's right node
    double pad;
};

struct RoughArgs {
    const RoughStep* steps;
    const double* denominators;   // [step][slot]: 1 + x_i dt (FP64 mode divides by it), +inf past the factors
    const double* reciprocals;    // [step][slot]: 1 / (1 + x_i dt) (mixed mode multiplies), 0 past the factors
    const uint32_t* sobol;        // [scramble][dim][33], dims = 2 * steps
    const uint32_t* store_nodes;  // ascending grid nodes whose spot is stored
    void* store;                  // [slot][path]: double (FP64 mode) or float
    uint64_t seed, units, total_paths;
    double spot, scale0;
    double v0, kappa, theta, nu, rho, rho_bar;
    double weights[kMaxRoughFactors];
    uint32_t n_steps, n_factors, n_store, rqmc, antithetic;
    uint32_t domain;              // Philox domain of the normals (Valuation 0, LsmTrain 1)
};

struct RoughProduct {    // terminal payoff read from one store slot
    int32_t kind;        // kEuropean or kDigital
    int32_t call, payout_asset;
    uint32_t slot;
    double strike, notional, payout, discount;
};

struct RoughItem {
    uint32_t product, scramble;
    uint64_t unit_begin, unit_end;
    uint32_t partial, pad;
};

struct RoughPayoffArgs {
    const RoughItem* items;
    const RoughProduct* products;
    const void* store;
    void* partials;
    uint64_t units, total_paths;
    uint32_t antithetic;
    uint32_t pad;
};

// ------------------------------------------ early exercise on lifted Heston --
// Regression on the features (spot, variance): standardized monomials of
// total degree <= 2 in the order 1, s, v, s^2, s v, v^2.
constexpr int kRlCols = 6;
constexpr int kRlGram = kRlCols * (kRlCols + 1) / 2 + kRlCols + 2;   // packed Gram, X'y, y'y, count
constexpr int kRlStats = 8;                                          // per-pass partial width (stats and variance)

struct RoughLsmDecision {
    double coef[kRlCols];          // per basis column, original order (0 for excluded columns)
    double diag[kRlCols];          // |R_pp| in pivot order
    double mean[2], variance[2], scale[2], inv_scale[2], zero_threshold[2];
    double rank_threshold, rss;
    uint32_t pivot[kRlCols];       // original column index per pivot position (active columns)
    uint32_t rank, continue_all, itm_rows, inactive;   // inactive: bit f set when feature f is inactive
    uint32_t active;               // number of active columns (the pivot entries in use)
    uint32_t pad;
};

// Per-date feature statistics of the in-the-money training rows.
struct RoughLsmScaling {
    double itm;                    // row count
    double mean[2], feature_scale[2];
    double variance[2], scale[2], zero_threshold[2];
    uint32_t inactive, pad;
};

struct RoughLsmArgs {
    RoughArgs train, value;        // the training and valuation engines on one grid and model
    const uint32_t* ex_nodes;      // grid node of each exercise date
    const double* ex_discount;     // DF(date)
    const double* ex_scale;        // physical spot per unit of F at the date's node
    double* store_s;               // training spots [date][path]
    double* store_v;               // training variances [date][path]
    double* cash;                  // training cash flow x DF(stopping date) [path]
    double* partials;              // [workgroup][kRlGram]
    RoughLsmScaling* scalings;     // [date]
    RoughLsmDecision* decisions;   // [non-terminal date]
    void* value_partials;          // BlockStat per valuation item
    unsigned long long* counts;    // valuation exercise counts [date]
    uint32_t* stops;               // optional: exercise-date index per valuation trajectory
    double* in_sample;             // mean of the training cash flows under the trained policy
    unsigned long long* error;
    uint64_t train_paths;          // training trajectories (units x scrambles x passes)
    uint64_t value_chunk;          // valuation units per item
    double strike, notional, itm_tol, abs_tol, rel_tol;
    uint32_t n_dates, call, columns;   // columns: 3 (degree 1) or 6 (degree 2)
    uint32_t n_wg;                 // workgroups of the training passes
    uint32_t value_items;
    uint32_t pad;
};

// Network continuation for the Bermudan under lifted Heston: one 2-64-64-1
// network per non-terminal date on (spot / K, variance), trained on the
// FP64 training stores of RoughLsmArgs.
struct RoughNnNorm {               // one date's row normalization
    float mu_s, inv_s, mu_v, inv_v, mu_y, sd_y, inv_y;
    uint32_t n;                    // in-the-money training rows
};

struct RoughNnArgs {
    const uint32_t* ex_nodes;      // as RoughLsmArgs
    const double* ex_discount;
    const double* ex_scale;
    const double* ex_times;        // year fraction of each exercise date
    const double* store_s;         // FP64 training stores [date][path] (shared with the polynomial policy)
    const double* store_v;
    double* cash;                  // the network's own training cash flows x DF(stop) [path]
    const float* init;             // initial parameters
    float* weights;                // [non-terminal date][kMlpStride]
    float* adam_m;                 // Adam moments [kMlpStride] (carried from date to date)
    float* adam_v;
    float4* rows;                  // (u_s, u_v, target, 1) per in-the-money training row
    float* grad;                   // [gradient workgroup][parameters + 1]
    double* stats;                 // [sweep workgroup][8]
    uint32_t* offsets;             // [sweep workgroup]: first row of the workgroup's rows
    RoughNnNorm* norms;            // [non-terminal date]
    double* in_sample;
    void* value_partials;          // BlockStat per valuation item (network policy)
    unsigned long long* counts;    // valuation exercise counts [date] (network policy)
    uint32_t* stops;               // optional: exercise-date index per valuation trajectory
    unsigned long long* error;
    uint64_t train_paths, value_chunk;
    double strike, notional, itm_tol, kappa, theta;
    float lr, beta1, beta2, eps;
    uint32_t n_dates, call, n_wg, batch, epochs, grad_wgs, rows_per_wg, value_items;
};

// ------------------------------------------------------- early exercise --
constexpr int kLsmThreads = 1024;     // induction workgroup: one per trade
constexpr int kLsmPassThreads = 256;  // weighted-Laguerre induction pass: one workgroup per chunk
constexpr uint32_t kLsmPassPaths = 8192;   // training paths per pass chunk
constexpr int kLsmMaxBasis = 4;       // spot feature, degree <= 3
constexpr int kLsmMaxDates = 1024;    // exercise dates per trade

enum LsmBasisKind : uint32_t {
    kLsmMonomial = 0,   // standardized monomials, CPQR pivot/rank rules, all-path decisions
    kLsmLaguerre = 1,   // exp(-x/2) L_k(x), x = S / K, in-the-money decisions
    kLsmMlp = 2,        // continuation network; decisions from its exercise region in moneyness
};

constexpr int kLsmMaxCuts = 64;   // the region search has 64 grid points: at most 63 sign changes

// Exercise region of the network policy at one (trade, date): in-the-money
// moneyness m = S / K is in the region iff `left` xor (number of cuts <= m
// is odd).  The cuts are the sign changes of immediate - continuation,
// located by bisection on the network itself.
struct LsmCuts {
    uint32_t count, left;
    float cut[kLsmMaxCuts];
};

// One exercise date of a trade: exact BS/Black-76 step and the escrowed
// dividend coordinate at the initial spot, S = a * S0 + b * F.
struct LsmDate {
    double time, sqrt_step, forward, discount;   // discount = DF(date)
    double a, b, ss;                             // post-dividend coordinate A, B and spot scale
    double m_half_tv;                            // -vol^2 t / 2
    double inv_discount;                         // 1 / DF(date) (mixed targets)
    float f_sqrt_step, f_forward, f_m_half_tv, f_pad;
};

// A sampling engine of a trade (training or valuation).
struct LsmEngine {
    uint64_t seed, units;         // pseudo units or RQMC points per scramble
    uint32_t scrambles, rqmc, antithetic, domain;
    uint32_t sobol_offset, pad;   // words into the batch Sobol table
};

struct LsmTrade {
    LsmEngine train, value;
    uint64_t store_offset;        // store elements before this trade ([date][path])
    uint64_t cash_offset;         // training paths before this trade
    uint64_t train_paths;         // units * scrambles * (1 + antithetic)
    uint32_t date_offset, n_dates;
    uint32_t decision_offset;     // non-terminal dates before this trade
    uint32_t count_offset;        // dates before this trade
    uint32_t call, zero_variance, single_terminal, basis;
    uint32_t columns;             // max degree + 1
    uint32_t pad;
    double strike, notional, s0, vol;
    double abs_tol, rel_tol, itm_tol, inv_strike;
    float f_vol, f_strike, f_s0, f_inv_strike;
    // fixed-policy Greeks (valuation kernel): Gamma bump, validation spot and vol bumps
    double gamma_bump, spot_bump, vol_bump;
    uint32_t risk, request_gamma;
    uint64_t stop_offset;         // valuation trajectories before this trade (stopping-index array)
};

constexpr int kLsmRiskComponents = 9;   // Delta .. Gamma difference (PE component indices 1..9)

// One date's fitted continuation model (ExerciseDecisionModel).
struct LsmDecision {
    double coef[kLsmMaxBasis];    // per basis column, original order
    double diag[kLsmMaxBasis];    // |R_pp| in pivot order
    double mean, variance, scale, inv_scale;   // monomial feature standardisation
    double rank_threshold, rss;
    double zero_threshold;        // FeatureScaling zero-scale threshold: 64 eps max(1, max |x|)
    uint32_t pivot[kLsmMaxBasis];
    uint32_t rank, continue_all, inactive, itm_rows;
    float fcoef[kLsmMaxBasis];    // FP32 copies for the mixed decisions
    float fmean, finv_scale;
};

struct LsmItem {
    uint32_t trade, scramble;
    uint64_t unit_begin, unit_end;
    uint32_t partial, pad;
};

struct LsmArgs {
    const LsmTrade* trades;
    const LsmDate* dates;
    const uint32_t* sobol;
    const LsmItem* items;
    void* store;                      // training spots, Real [date][path] per trade
    void* cash;                       // training cash flow x DF(stopping date), Real
    LsmDecision* decisions;
    unsigned long long* counts;       // valuation exercise counts [trade][date]
    double* in_sample;                // per trade: mean of the training cash flows
    void* partials;                   // BlockStat per valuation item
    void* risk_partials;              // kLsmRiskComponents BlockStats per valuation item (fixed-policy Greeks)
    unsigned long long* error;
    const LsmCuts* cuts;              // network policy: [trade][non-terminal date]
    uint32_t* stops;                  // optional: exercise-date index per valuation trajectory
    uint32_t n_items, n_trades;
    double* gram;                     // weighted-Laguerre passes: 16 partial sums per (trade, chunk)
    uint32_t gram_chunks;             // chunks per trade in `gram` (the largest trade's)
    uint32_t max_dates;               // most exercise dates of any trade
};

// Andersen-Broadie bounds of one trained American trade (the valuation
// engine supplies the outer paths).
struct LsmDualArgs {
    LsmArgs lsm;            // trained trade: trades, dates, decisions (one trade)
    double* cont;           // [trajectory][non-terminal date] inner rollout means
    void* partials;         // BlockStat [block][3]: lower bound, upper bound, gap
    uint64_t inner_seed;
    uint64_t stride;        // max(N2, N3): inner path index stride per trajectory
    uint32_t n2, n3;        // continuation / exercise inner paths
    uint32_t lanes;         // 2 with antithetic outer paths
    uint32_t pad;
};

// The learned policy's persistent training launch: one network per exercise
// date shared by every trade of the launch (inputs moneyness and volatility,
// output the continuation value per unit strike), trained on the pooled
// in-the-money training rows.
struct MlpLsmArgs {
    LsmArgs lsm;
    const float* init;         // initial parameters (every trade)
    float* weights;            // FP32 master parameters [trade][stride]
    float* adam_m;             // Adam moments [trade][stride]
    float* adam_v;
    float4* rows;              // training rows (u1, u2, target, 1), one per training path of each trade
    LsmCuts* cuts;             // [trade][non-terminal date]
    unsigned int* error;       // bit 0: an exercise region needed more than kLsmMaxCuts cuts
    unsigned long long* rows_trained;   // [trade]: rows run through the network (epochs x rows, all dates)
    size_t cash_bytes;         // bytes of the training cash-flow vector (copied for the polynomial judge)
    uint32_t n_trades, batch, epochs, pad;
    float lr, beta1, beta2, eps;
    float grid_lo, grid_hi;    // moneyness range of the exercise-region search
};

// Kernel arguments of the local-vol Log-Euler kernel.
// Forward normalizers of the local-vol runtime recompiled at spot -/+ the
// validation bump, per step (pathwise_local_vol.rs:24-69).
struct LvBumpStep {
    double forward_dn, ratio_dn, forward_up, ratio_up;
};

enum LvRiskFlags : uint32_t { kLvRiskBumps = 1u, kLvRiskVega = 2u, kLvRiskGamma = 4u };

struct LvArgs {
    const LvStep* steps;
    const BridgeOp* program;
    const LvOp32* ops32;       // fused program (bridge pushes + steps), mixed kernel
    const LvOp64* ops64;       // same program with FP64 constants
    uint32_t n_ops;
    uint32_t ops_stride;       // records per scramble copy of ops32 (0: one shared copy)
    ProductHot hot[kMaxProducts];
    // Uniform bins over [x0, xl] (mixed kernel): the FP32 coordinate xx maps
    // to bin trunc(fma(xx, inv_bin_w32, bin_off32)); bin_cell[b] counts the
    // interior nodes whose own bin is below b, and no two interior nodes
    // share a bin.
    const uint16_t* bin_cell;
    uint32_t bins;
    float inv_bin_w32, bin_off32, x0_32, xl_32;
    const double* grid;        // [nt][nx] local variance (FP64)
    const float* grid32;       // FP32 copy for the mixed kernel
    const double* x_nodes;
    const float* x_nodes32;
    const uint32_t* sobol;
    void* partials;
    ProductDesc products[kMaxProducts];
    double weights[kMaxProducts];  // unused; per-node weights live in node_weight
    const double* node_weight;     // [node][product]
    uint64_t seed;
    uint64_t units;
    double spot;
    uint32_t n_steps;
    uint32_t n_program;
    uint32_t nt, nx;
    uint32_t n_products;
    uint32_t rqmc;
    uint32_t antithetic;
    uint32_t bridge;
    uint32_t depth;
    uint32_t sobol_dims;
    uint32_t node0_mask;       // products with a fixing at the valuation node
    // a graph product (one product per launch): its register program, one
    // segment per observation, constants, output operand, registers in LDS
    const GraphOp* graph_ops;
    const uint32_t* graph_segs;
    const double* graph_consts;
    uint32_t graph_output, graph_regs, graph_lds_offset, graph_pad;
    // risk kernel only (one plan): bumped runtimes and the node-tangent table
    const LvBumpStep* bump_steps;
    double spot_dn, spot_up, spot_bump;
    double scale_dn, scale_up;     // bumped spot / spot (mixed mode)
    uint32_t risk;                 // LvRiskFlags
    uint32_t risk_lds_offset;      // byte offset of the 2 sqrt(v) table in LDS
};

}  // namespace pe
