// Host-side plan of the engine: FP64 plan compilation mirroring
// SimulationPlan::compile (crates/pricing/src/engine/compile/simulation.rs:78-640,
// crates/pricing/src/engine/compile/local_vol.rs:14-103), flattened into the
// POD tables the kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "pe.h"
#include "pe_types.hpp"

namespace pe {

struct PlanError : std::runtime_error {
    pe_status status;
    PlanError(pe_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

// The calling thread's stream of the current entry-point call (nullptr
// outside one: the legacy stream).
hipStream_t current_stream();

// Device buffer owned by a plan.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void reserve(size_t n) {
        if (n <= bytes) return;
        release();
        if (hipMalloc(&ptr, n) != hipSuccess) {
            ptr = nullptr;
            throw PlanError(PE_ERR_ALLOCATION, "hipMalloc failed");
        }
        bytes = n;
    }
    template <typename T>
    void upload(const std::vector<T>& v) {
        reserve(v.size() * sizeof(T) + 16);
        if (v.empty()) return;
        hipStream_t s = current_stream();
        if (hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            throw PlanError(PE_ERR_DEVICE, "hipMemcpy to device failed");
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Log-linear discount curve (crates/pricing/src/market/curve.rs:39-212).
struct HostCurve {
    std::vector<double> times, log_dfs;
    void init(const pe_curve& c);
    double log_discount(double t) const;
    pe_curve_region region(double t) const;
};

// Launch state of a local-vol evaluation over one set of plans, cached on the
// set's first plan so repeated evaluations only launch and read back.
struct LvLaunch {
    std::vector<uint64_t> key;       // ids of the plans, in order
    int backend = 0;
    LvArgs args{};
    dim3 grid;
    size_t lds = 0;
    DeviceBuffer d_steps, d_node_weight, d_ops32, d_ops64, d_bins, d_bumps, partials, err, d_graph_ops, d_graph_segs,
        d_graph_consts;
    bool risk = false;               // one plan with Delta/Vega/Gamma: lv_risk_kernel
    bool graph = false;              // one graph product: the graph flavour of lv_kernel
    void* host_partials = nullptr;   // pinned
    int* host_err = nullptr;         // pinned
    size_t partial_count = 0;
    ~LvLaunch() {
        if (host_partials) (void)hipHostFree(host_partials);
        if (host_err) (void)hipHostFree(host_err);
    }
};

// Escrowed affine discrete dividends (crates/pricing/src/market/dividend.rs:96-547):
// S = A*S0 + B*F, cash funded by a reserve carried on the market curves.
struct HostCoord {
    double a = 0.0, b = 1.0, ss = 1.0;
};

struct HostDividends {
    struct Event {
        double ex_time, cash, alpha, beta;
    };
    struct Entry {
        double ex_time;
        HostCoord before, after;
    };
    double spot = 0.0, reserve = 0.0;
    std::vector<Event> events;
    std::vector<Entry> timeline;
    const HostCurve* discount = nullptr;
    const HostCurve* dividend = nullptr;

    bool empty() const { return events.empty(); }
    void compile(double s0, const pe_dividend_quote* q, uint32_t n, const HostCurve* p, const HostCurve* d);
    HostCoord initial() const;
    HostCoord carry(HostCoord c, double start, double end) const;
    HostCoord after_time(double t) const;
};

// Lifted Heston model of a rough-volatility plan.
struct HostLifted {
    double v0 = 0, kappa = 0, theta = 0, nu = 0, rho = 0;
    std::vector<double> weights, rates;
};

struct BsLaunch;
struct BsLaunchDeleter {
    void operator()(BsLaunch* l) const;
};
struct RoughLaunch;
struct RoughLaunchDeleter {
    void operator()(RoughLaunch* l) const;
};
struct LsmLaunch;
struct LsmLaunchDeleter {
    void operator()(LsmLaunch* l) const;
};
struct RoughLsmLaunch;
struct RoughLsmLaunchDeleter {
    void operator()(RoughLsmLaunch* l) const;
};

// A graph product's tape for the BS batch kernel (compile_graph, pe_graph.cpp).
struct GraphTape {
    std::vector<GraphOp> ops;        // scheduled operations
    std::vector<uint32_t> segs;      // days.size() + 2 boundaries: stage s (-1 = before the path,
                                     // i = after observation i) is ops [segs[s + 1], segs[s + 2])
    std::vector<double> consts;      // folded constants (operand kGraphConst | index)
    std::vector<int32_t> days;       // sorted unique observation days of the tape
    std::vector<int32_t> pre_days;   // days read before their dividend
    int32_t last_payment = 0;        // discounting date of the single output
    int32_t expiry_day = 0;          // last observation (or payment without observations)
    uint16_t output = 0;             // register or constant operand of the value
    uint16_t regs = 1;               // registers in use
};

// Validates and compiles `pr` (kind PE_PRODUCT_GRAPH); `discount(day)` is the
// discount factor of a payment day.  Throws PlanError.
void compile_graph(const pe_product& pr, const std::function<double(int32_t)>& discount, GraphTape& g);

struct HostPlan {
    uint64_t id = 0;                 // unique per compiled plan, never reused
    std::mutex mu;                   // held by an evaluation that uses this plan's caches
    std::unique_ptr<LvLaunch> lv_launch;
    std::unique_ptr<BsLaunch, BsLaunchDeleter> bs_launch;
    std::unique_ptr<RoughLaunch, RoughLaunchDeleter> rough_launch;
    pe_request req{};
    std::vector<int32_t> dates;
    std::vector<double> weights, fixings;

    HostCurve discount_curve, dividend_curve;
    HostDividends dividends;
    double spot = 0.0;
    std::vector<int32_t> obs_days;        // kNoDay: dividend-only bridge point
    std::vector<double> obs_times, obs_forwards;
    std::vector<HostCoord> obs_coord, obs_pre;
    std::vector<uint32_t> obs_flags;      // BsObsFlags
    bool continuous = false, resolved = false, single_terminal = false;
    double resolved_survival = 1.0;
    int32_t expiry_obs = -1;
    double expiry = 0.0, forward = 0.0, discount = 1.0;
    double vol = 0.0, total_variance = 0.0;
    double gamma_bump = 0.0, spot_bump = 0.0, vol_bump = 0.0;
    bool risk = false, request_gamma = false, request_delta = false, request_vega = false;
    uint64_t block_size = 4096;
    pe_mc_diagnostics diag{};           // MonteCarloDiagnostics reported by every evaluation

    // local vol
    bool lv = false;
    std::vector<double> lv_t, lv_x, lv_v;
    double lv_cap = 0.0;
    std::vector<double> nodes;          // simulation grid
    std::vector<double> node_forwards;
    std::vector<uint32_t> obs_nodes;    // node index per observation
    std::vector<LvStep> steps;
    std::vector<LvBumpStep> lv_bumps;   // risk: runtimes at spot -/+ spot_bump, per step
    std::vector<BridgeOp> program;
    uint32_t bridge_depth = 0;

    ProductDesc product{};
    bool graph = false;                 // graph product: the tape below is the payoff
    GraphTape graph_tape;
    std::vector<pe_graph_node> graph_nodes;
    std::vector<pe_graph_cashflow> graph_cashflows;
    std::vector<uint32_t> graph_underlyings;
    std::vector<double> obs_weight;     // per observation (Asian weight or 1 for fixings)

    // RQMC
    uint32_t sobol_dims = 0;
    std::vector<uint32_t> host_sobol;   // scrambled direction table, built on first use (allocation-free path)

    // rough-volatility adapter (lifted Heston); the carrier's BS fields above
    // supply the payoff, discount and observation times
    bool rough = false;
    HostLifted lifted;
    double rough_max_step = 0.0;
    std::vector<double> rough_events;
    std::unique_ptr<RoughLsmLaunch, RoughLsmLaunchDeleter> rough_lsm_launch;   // Bermudan under lifted Heston

    // BS/Black-76 trade descriptor and observation table for the batch kernel
    BsTrade trade{};
    std::vector<BsObs> bs_obs;
    std::vector<BsObs32> bs_obs32;        // FP32 constants of the mixed batch kernel

    // early exercise (American product): exercise dates and the LSM runtime
    bool american = false;
    LsmTrade lsm_trade{};
    std::vector<LsmDate> lsm_dates;
    std::vector<int32_t> lsm_collisions;  // per exercise date: an ex-dividend time equals it
    pe_mc_diagnostics train_diag{};       // training engine's RQMC checksums
    std::unique_ptr<LsmLaunch, LsmLaunchDeleter> lsm_launch;

    // device tables
    DeviceBuffer d_obs, d_sobol, d_steps, d_program, d_grid, d_grid32, d_x, d_x32, d_node_weight;

    double forward_at(double t) const { return forward_at_spot(t, spot); }
    double forward_at_spot(double t, double s) const;
};

void compile_host_plan(HostPlan& p, const pe_request& r, const pe_exec_policy* policy);
void compile_diagnostics(HostPlan& p, const pe_exec_policy* policy, uint32_t sobol_dims);
void rqmc_checksums(uint64_t seed, uint32_t scrambles, uint32_t dims, uint8_t (&direction)[32],
                    uint8_t (&scramble)[32]);
void compile_rough_plan(HostPlan& p, const pe_request& carrier, const pe_lifted_heston& m, double maximum_step,
                        const pe_exec_policy* policy);
void lifted_from_rough(double hurst, uint32_t factors, double ratio, std::vector<double>& weights,
                       std::vector<double>& rates);
// LocalVolTimeGrid::compile: sorted unique events, ceil(gap / maximum_step) equal substeps.
std::vector<double> compile_time_grid(std::vector<double> events, double maximum_step);
std::vector<uint32_t> build_sobol_table(uint64_t seed, uint32_t scrambles, uint32_t dims);
void compile_bridge_program(const std::vector<double>& times, std::vector<BridgeOp>& program,
                            std::vector<double>& inv_sqrt_dt, uint32_t& depth);
const uint32_t* direction_words(uint32_t* dimensions);  // embedded Joe--Kuo asset (host order)

constexpr int32_t kNoDay = INT32_MIN;

// BLAKE3-256 hasher (pe_blake3.cpp), unkeyed, incremental.
class Blake3 {
public:
    Blake3();
    void reset();
    void update(const void* data, size_t n);
    void finalize(uint8_t out[32]) const;

private:
    void compress_buffer();
    void finish_chunk();
    uint32_t cv_[8];
    uint8_t block_[64];
    uint64_t chunk_counter_;
    uint32_t block_len_, blocks_compressed_;
    uint32_t stack_[54][8];
    uint32_t stack_len_;
};

// Fixed device workspace of the allocation-free entry point (pe_noalloc.hip).
size_t scratch_bytes();
hipError_t scratch_setup(int device);   // with `device` current; once per device
void* scratch_pointer(int device);

// Launchers (pe_kernels_f64.hip / pe_kernels_f32.hip).
hipError_t launch_bs_batch_f64(const BsBatchArgs& a, bool risk, bool cont, int kind, uint32_t graph_regs,
                               hipStream_t s);
hipError_t launch_bs_batch_f32(const BsBatchArgs& a, bool risk, bool cont, int kind, uint32_t graph_regs,
                               hipStream_t s);
// The same launch as a kernel node of `graph` (no dependencies).
hipError_t add_bs_batch_node_f64(hipGraph_t graph, const BsBatchArgs& a, bool risk, bool cont, int kind,
                                 uint32_t graph_regs);
hipError_t add_bs_batch_node_f32(hipGraph_t graph, const BsBatchArgs& a, bool risk, bool cont, int kind,
                                 uint32_t graph_regs);
hipError_t launch_rough_f64(const RoughArgs& a, const RoughPayoffArgs& pa, dim3 grid, uint32_t payoff_items,
                            unsigned long long* error, hipStream_t s);
hipError_t launch_rough_f32(const RoughArgs& a, const RoughPayoffArgs& pa, dim3 grid, uint32_t payoff_items,
                            unsigned long long* error, hipStream_t s);
hipError_t launch_rough_lsm_f64(const RoughLsmArgs& a, dim3 train_grid, dim3 value_grid, hipStream_t s,
                               hipEvent_t paths_done, hipEvent_t between, void* cash_copy);
hipError_t launch_rough_lsm_f32(const RoughLsmArgs& a, dim3 train_grid, dim3 value_grid, hipStream_t s,
                               hipEvent_t paths_done, hipEvent_t between, void* cash_copy);
hipError_t launch_rough_nn_value_f64(const RoughNnArgs& A, const RoughArgs& a, dim3 grid, hipStream_t s);
hipError_t launch_rough_nn_value_f32(const RoughNnArgs& A, const RoughArgs& a, dim3 grid, hipStream_t s);
hipError_t launch_rough_nn_date(const RoughNnArgs& A, uint32_t d, hipStream_t s);
hipError_t launch_rough_nn_train(const RoughNnArgs& A, uint32_t d, uint32_t steps, uint32_t adam_t, hipStream_t s);
hipError_t launch_rough_nn_final(const RoughNnArgs& A, hipStream_t s);
hipError_t launch_lsm_dual_f64(const LsmDualArgs& da, uint64_t pairs, uint32_t outer_blocks, uint32_t basis,
                               hipStream_t s);
hipError_t launch_lv_risk_f64(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s);
hipError_t launch_lv_risk_f32(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s);
hipError_t launch_lsm_f64(const LsmArgs& train, uint32_t train_items, const LsmArgs& value, uint32_t value_items,
                          uint32_t basis, bool risk, hipStream_t s, hipEvent_t paths_done, hipEvent_t between);
hipError_t launch_lsm_value_f32(const LsmArgs& value, uint32_t value_items, uint32_t basis, bool risk, hipStream_t s);
hipError_t launch_lsm_f32(const LsmArgs& train, uint32_t train_items, const LsmArgs& value, uint32_t value_items,
                          uint32_t basis, bool risk, hipStream_t s, hipEvent_t paths_done, hipEvent_t between);
hipError_t launch_lsm_mlp_f64(const LsmArgs& train, uint32_t train_items, const MlpLsmArgs& m, const LsmArgs& judge,
                              const LsmArgs& value, const LsmArgs& judge_value, uint32_t value_items, bool risk,
                              hipStream_t s, const hipEvent_t* ev);
hipError_t launch_lsm_mlp_f32(const LsmArgs& train, uint32_t train_items, const MlpLsmArgs& m, const LsmArgs& judge,
                              const LsmArgs& value, const LsmArgs& judge_value, uint32_t value_items, bool risk,
                              hipStream_t s, const hipEvent_t* ev);
hipError_t launch_mlp_gradient(const float* params, const float* x, const float* y, uint32_t rows, uint32_t rows_wg,
                               uint32_t wgs, float* partials, hipStream_t s);
hipError_t launch_mlp_forward(const float* params, const float* x, uint32_t rows, float* out, hipStream_t s);
constexpr uint32_t kMlpParameterCount = 4417;
hipError_t launch_normal_mixed_error(uint64_t first, uint64_t count, uint32_t blocks, double* block_err,
                                     uint32_t* block_word, hipStream_t s);
hipError_t launch_lv_f64(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s);
hipError_t launch_lv_f32(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s);
hipError_t launch_lv_graph_f64(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s);
hipError_t launch_lv_graph_f32(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s);
hipError_t launch_dupire(const double* theta_t, const double* theta_v, const double* theta_d, uint32_t nk,
                         double slope, double rho, double eta, double gamma, const double* t, uint32_t nt,
                         const double* x, uint32_t nx, double floor, double cap, double* out, uint32_t* repairs,
                         int* status, hipStream_t s);

}  // namespace pe
