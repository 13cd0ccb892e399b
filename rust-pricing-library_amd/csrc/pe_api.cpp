// C ABI of the engine (include/pe.h): plan lifecycle, launches, block-partial
// merging and result assembly.
//
// Result assembly mirrors execute_pseudo / execute_rqmc and
// estimate_from_statistics (crates/pricing/src/engine/risk/valuation.rs:126-437,
// crates/pricing/src/engine/risk/report.rs:454-473): pseudo-MC statistics over
// sampling units, RQMC statistics as a two-pass over per-scramble means.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pe.h"
#include "pe_device.hpp"
#include "pe_plan.hpp"

struct pe_plan {
    pe::HostPlan host;
};

namespace pe {
namespace {

constexpr double kNormal95 = 1.959963984540054;  // crates/pricing/src/engine/plan/simulation.rs:24

thread_local std::string g_error;
thread_local double g_kernel_ms = 0.0;
thread_local uint32_t g_kernel_launches = 0;

std::mutex g_init_mutex;
bool g_initialised = false;
int g_device = 0;

// One HIP stream per entry-point call: a call takes a stream from the pool
// (or creates one), makes it this thread's current stream, and returns it
// when it ends, so concurrent calls from several host threads never share a
// stream or a graph capture.
std::mutex g_pool_mutex;
std::vector<hipStream_t> g_free_streams;
thread_local hipStream_t t_stream = nullptr;

// A synchronous copy ordered on the call's stream (no legacy-stream
// synchronisation with other threads' work).
hipError_t copy_on_stream(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, t_stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(t_stream);
}

struct CallStream {
    hipStream_t prev = nullptr, own = nullptr;
    CallStream() {
        {
            std::lock_guard<std::mutex> lock(g_pool_mutex);
            if (!g_free_streams.empty()) {
                own = g_free_streams.back();
                g_free_streams.pop_back();
            }
        }
        if (!own && hipStreamCreateWithFlags(&own, hipStreamNonBlocking) != hipSuccess)
            throw PlanError(PE_ERR_DEVICE, "hipStreamCreate failed");
        prev = t_stream;
        t_stream = own;
    }
    ~CallStream() {
        t_stream = prev;
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        g_free_streams.push_back(own);
    }
    CallStream(const CallStream&) = delete;
    CallStream& operator=(const CallStream&) = delete;
};

// Locks the plans of one call (each plan's launch caches and device tables)
// in address order, so calls on overlapping plan sets cannot deadlock.
struct PlanLocks {
    std::vector<std::unique_lock<std::mutex>> held;
    explicit PlanLocks(std::vector<HostPlan*> ps) {
        std::sort(ps.begin(), ps.end());
        ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
        for (HostPlan* q : ps) held.emplace_back(q->mu);
    }
};

pe_status fail(pe_status s, const std::string& what) {
    g_error = what;
    return s;
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw PlanError(PE_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

struct Stat {  // Neumaier sum + centred moments (reduction.rs:3-161)
    double sum = 0.0, corr = 0.0, mean = 0.0, m2 = 0.0;
    uint64_t n = 0;

    void add_sum(double v) {
        const double t = sum + v;
        if (std::fabs(sum) >= std::fabs(v)) corr += (sum - t) + v; else corr += (v - t) + sum;
        sum = t;
    }
    void merge(const BlockStat& b) {
        if (b.n == 0) return;
        add_sum(b.sum);
        add_sum(b.corr);
        if (n == 0) {
            mean = b.mean; m2 = b.m2; n = b.n;
            return;
        }
        const uint64_t nn = n + b.n;
        const double delta = b.mean - mean;
        const double lw = static_cast<double>(n), rw = static_cast<double>(b.n), tw = static_cast<double>(nn);
        mean = mean + delta * (rw / tw);
        m2 = (m2 + b.m2) + delta * delta * (lw * rw / tw);
        n = nn;
    }
    double total() const { return sum + corr; }
    static Stat two_pass(const std::vector<double>& v) {
        Stat s;
        for (double x : v) s.add_sum(x);
        s.n = v.size();
        if (v.empty()) return s;
        s.mean = s.total() / static_cast<double>(v.size());
        Stat q;
        for (double x : v) {
            const double d = x - s.mean;
            q.add_sum(d * d);
        }
        s.m2 = q.total();
        return s;
    }
};

pe_estimate estimate(const Stat& s, uint64_t units) {
    const double inv = 1.0 / static_cast<double>(units);
    const double value = s.total() * inv;
    const double var = s.n > 1 ? s.m2 / static_cast<double>(s.n - 1) : 0.0;
    const double se = std::sqrt(var * inv);
    return pe_estimate{value, se, value - kNormal95 * se, value + kNormal95 * se, units};
}

void ensure_init() {
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (g_initialised) return;
    int count = 0;
    check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (count <= 0) throw PlanError(PE_ERR_DEVICE, "no HIP device");
    check(hipSetDevice(g_device), "hipSetDevice");
    check(scratch_setup(g_device), "workspace symbol address");
    g_initialised = true;
}

// Converts block partials to per-component statistics and fills the result.
void assemble(const HostPlan& p, const BlockStat* partials, uint32_t blocks_x, uint32_t scrambles,
              uint32_t comps, uint32_t comp_offset, uint32_t result_comps, pe_result* out) {
    const pe_engine& e = p.req.engine;
    const bool rqmc = e.kind == PE_ENGINE_RQMC;
    const uint64_t independent = rqmc ? e.scrambles : e.units;
    std::vector<Stat> stats(PE_COMPONENTS);
    if (!rqmc) {
        for (uint32_t c = 0; c < result_comps; ++c)
            for (uint32_t b = 0; b < blocks_x; ++b) stats[c].merge(partials[static_cast<size_t>(b) * comps + comp_offset + c]);
    } else {
        // only each scramble's compensated total is used: the blocks' sums
        // are added in the same order as Stat::merge adds them, without its
        // centred moments
        std::vector<double> means(scrambles);
        for (uint32_t c = 0; c < result_comps; ++c) {
            for (uint32_t s = 0; s < scrambles; ++s) {
                Stat within;
                const BlockStat* row = partials + static_cast<size_t>(s) * blocks_x * comps + comp_offset + c;
                for (uint32_t b = 0; b < blocks_x; ++b) {
                    const BlockStat& bs = row[static_cast<size_t>(b) * comps];
                    if (bs.n == 0) continue;
                    within.add_sum(bs.sum);
                    within.add_sum(bs.corr);
                }
                means[s] = within.total() / static_cast<double>(e.units);
            }
            stats[c] = Stat::two_pass(means);
        }
    }
    for (int c = 0; c < PE_COMPONENTS; ++c) {
        out->estimates[c] = estimate(stats[c], independent);
        out->component_sum[c] = stats[c].total();
        out->component_mean[c] = stats[c].mean;
        out->component_m2[c] = stats[c].m2;
    }
    double sv = 0.0;
    if (p.total_variance != 0.0) {
        if (stats[0].n < 2) throw PlanError(PE_ERR_INSUFFICIENT_UNITS, "insufficient sampling units");
        sv = stats[0].m2 / static_cast<double>(stats[0].n - 1);
    }
    out->sampling_variance = sv;
    out->estimator_variance = sv / static_cast<double>(independent);
    out->independent_units = independent;
    const uint64_t mult = e.antithetic ? 2 : 1;
    out->evaluated_paths = e.units * (rqmc ? e.scrambles : 1) * mult;
    out->time_steps = p.lv ? static_cast<uint32_t>(p.nodes.size() - 1) : static_cast<uint32_t>(p.obs_times.size());
    out->path_steps = out->evaluated_paths * out->time_steps;
    out->sobol_dimension = rqmc ? p.sobol_dims : 0;
    out->diagnostics = p.diag;
}

void assemble(const HostPlan& p, const std::vector<BlockStat>& partials, uint32_t blocks_x, uint32_t scrambles,
              uint32_t comps, uint32_t comp_offset, uint32_t result_comps, pe_result* out) {
    assemble(p, partials.data(), blocks_x, scrambles, comps, comp_offset, result_comps, out);
}

dim3 grid_for(const pe_engine& e) {
    const uint64_t blocks = (e.units + kBlockThreads - 1) / kBlockThreads;
    const uint32_t bx = static_cast<uint32_t>(std::min<uint64_t>(blocks, 1u << 16));
    return dim3(bx, e.kind == PE_ENGINE_RQMC ? e.scrambles : 1u, 1u);
}

void ensure_sobol(HostPlan& p) {
    const pe_engine& e = p.req.engine;
    if (e.kind != PE_ENGINE_RQMC || p.d_sobol.ptr) return;
    p.d_sobol.upload(build_sobol_table(e.seed, e.scrambles, p.sobol_dims));
}

// The two timing events of this thread's calls, created once.
struct CallEvents {
    hipEvent_t start = nullptr, stop = nullptr;
    CallEvents() {
        check(hipEventCreate(&start), "hipEventCreate");
        check(hipEventCreate(&stop), "hipEventCreate");
    }
    ~CallEvents() {
        if (start) (void)hipEventDestroy(start);
        if (stop) (void)hipEventDestroy(stop);
    }
};

CallEvents& call_events() {
    thread_local CallEvents ev;
    return ev;
}

template <typename F>
float timed(F&& launch) {
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    check(hipEventRecord(a, t_stream), "hipEventRecord");
    launch();
    check(hipEventRecord(b, t_stream), "hipEventRecord");
    check(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

// ---------------------------------------------------------- BS batches --
// A batch of BS/Black-76 plans runs as one launch per kernel flavour
// (price-only / risk x discrete / continuous barrier), one workgroup per work
// item.  Work items are 4096-unit ranges of one trade and scramble, ordered
// longest-first; each has a fixed partial slot so the per-trade merge below
// is in unit order whatever the schedule.
constexpr uint64_t kItemUnits = 4096;
// groups: bit 0 risk, bit 1 continuous barrier; group 4: graph products,
// group 5: graph products with Greeks
// Launch groups: fixed products with discrete monitoring by product family
// (European, Digital, Barrier, Asian, Lookback) and price / risk, so each
// launch runs one family's payoff code; continuous barriers; graph products.
constexpr int kBsFamilies = 5;
constexpr int kBsContGroup = 2 * kBsFamilies;
constexpr int kBsGraphGroup = kBsContGroup + 2;
constexpr int kBsGroups = kBsGraphGroup + 2;

int bs_group_of(const HostPlan& p) {
    const int risk = p.risk ? 1 : 0;
    if (p.graph) return kBsGraphGroup + risk;
    if (p.continuous) return kBsContGroup + risk;
    const int kind = p.product.kind;
    if (kind < 0 || kind >= kBsFamilies) throw PlanError(PE_ERR_INVALID_ARGUMENT, "product family of a batch trade");
    return 2 * kind + risk;
}
bool bs_group_risk(int g) { return (g & 1) != 0; }
bool bs_group_cont(int g) { return g == kBsContGroup || g == kBsContGroup + 1; }
// the product family compiled into the group's kernel (-1: any)
int bs_group_kind(int g) { return g < kBsContGroup ? g / 2 : -1; }

}  // namespace

struct BsGroup {
    bool risk = false, cont = false;
    uint32_t graph_regs = 0;        // graph group: registers of its largest tape
    uint32_t components = 1;
    std::vector<BsItem> items;
    size_t partial_base = 0;        // BlockStats before this group's slots
    DeviceBuffer d_items;
};

struct BsPlanSlot {
    int group = 0;
    uint32_t first_partial = 0;     // within the group
    uint32_t chunks = 0;
};

struct BsLaunch {
    std::vector<uint64_t> key;
    int backend = 0;
    std::vector<BsTrade> trades;
    std::vector<BsPlanSlot> slots;
    BsGroup groups[kBsGroups];
    DeviceBuffer d_trades, d_obs, d_obs32, d_sobol, partials, err, d_graph_ops, d_graph_segs, d_graph_consts;
    std::vector<BlockStat> host_partials;
    size_t partial_count = 0;
    // persistent grid size per group, chosen by the tuner on the first
    // evaluation; the tuned launches are then captured in one HIP graph
    uint32_t grid[kBsGroups] = {};
    bool tuned = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    ~BsLaunch() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

void BsLaunchDeleter::operator()(BsLaunch* l) const { delete l; }

namespace {

BsLaunch& prepare_bs(std::vector<HostPlan*>& plans, pe_backend backend) {
    HostPlan& p0 = *plans[0];
    std::vector<uint64_t> key;
    for (const HostPlan* q : plans) key.push_back(q->id);
    if (p0.bs_launch && p0.bs_launch->key == key && p0.bs_launch->backend == static_cast<int>(backend))
        return *p0.bs_launch;
    auto L = std::make_unique<BsLaunch>();
    L->key = key;
    L->backend = static_cast<int>(backend);
    std::vector<BsObs> obs;
    std::vector<BsObs32> obs32;
    std::vector<uint32_t> sobol;
    std::vector<GraphOp> graph_ops;
    std::vector<uint32_t> graph_segs;
    std::vector<double> graph_consts;
    for (int g = 0; g < kBsGroups; ++g) {
        L->groups[g].risk = bs_group_risk(g);
        L->groups[g].cont = bs_group_cont(g);
        L->groups[g].components = L->groups[g].cont ? PE_COMPONENTS : (L->groups[g].risk ? 10u : 1u);
    }
    uint32_t group_partials[kBsGroups] = {};
    for (size_t ti = 0; ti < plans.size(); ++ti) {
        HostPlan& p = *plans[ti];
        if (p.lv) throw PlanError(PE_ERR_INVALID_ARGUMENT, "local-vol plans are evaluated by pe_plans_evaluate_shared");
        BsTrade t = p.trade;
        t.obs_offset = static_cast<uint32_t>(obs.size());
        obs.insert(obs.end(), p.bs_obs.begin(), p.bs_obs.end());
        obs32.insert(obs32.end(), p.bs_obs32.begin(), p.bs_obs32.end());
        const pe_engine& e = p.req.engine;
        const uint32_t scrambles = e.kind == PE_ENGINE_RQMC ? e.scrambles : 1u;
        if (e.kind == PE_ENGINE_RQMC) {
            t.sobol_offset = static_cast<uint32_t>(sobol.size());
            const std::vector<uint32_t> tab = build_sobol_table(e.seed, e.scrambles, p.sobol_dims);
            if (sobol.size() + tab.size() > 0xffffffffull) throw PlanError(PE_ERR_SIZE_LIMIT, "Sobol tables");
            sobol.insert(sobol.end(), tab.begin(), tab.end());
        }
        if (p.graph) {
            const GraphTape& gt = p.graph_tape;
            t.graph_ops = static_cast<uint32_t>(graph_ops.size());
            t.graph_segs = static_cast<uint32_t>(graph_segs.size());
            t.graph_consts = static_cast<uint32_t>(graph_consts.size());
            graph_ops.insert(graph_ops.end(), gt.ops.begin(), gt.ops.end());
            graph_segs.insert(graph_segs.end(), gt.segs.begin(), gt.segs.end());
            graph_consts.insert(graph_consts.end(), gt.consts.begin(), gt.consts.end());
            if (gt.segs.size() != p.bs_obs.size() + 2) throw PlanError(PE_ERR_INVALID_ARGUMENT, "graph schedule");
            uint32_t& regs = L->groups[bs_group_of(p)].graph_regs;
            regs = std::max<uint32_t>(regs, gt.regs);
        }
        L->trades.push_back(t);
        const int g = bs_group_of(p);
        BsGroup& grp = L->groups[g];
        BsPlanSlot slot;
        slot.group = g;
        slot.chunks = static_cast<uint32_t>((e.units + kItemUnits - 1) / kItemUnits);
        slot.first_partial = group_partials[g];
        for (uint32_t sc = 0; sc < scrambles; ++sc)
            for (uint32_t c = 0; c < slot.chunks; ++c) {
                BsItem it{};
                it.trade = static_cast<uint32_t>(ti);
                it.scramble = sc;
                it.unit_begin = static_cast<uint64_t>(c) * kItemUnits;
                it.unit_end = std::min<uint64_t>(it.unit_begin + kItemUnits, e.units);
                it.partial = slot.first_partial + sc * slot.chunks + c;
                grp.items.push_back(it);
            }
        group_partials[g] += scrambles * slot.chunks;
        L->slots.push_back(slot);
    }
    size_t base = 0;
    for (int g = 0; g < kBsGroups; ++g) {
        BsGroup& grp = L->groups[g];
        grp.partial_base = base;
        base += static_cast<size_t>(group_partials[g]) * grp.components;
        // longest items first: the work queue then ends with short items
        std::stable_sort(grp.items.begin(), grp.items.end(), [&](const BsItem& x, const BsItem& y) {
            const BsTrade& tx = L->trades[x.trade];
            const BsTrade& ty = L->trades[y.trade];
            const uint64_t cx = (x.unit_end - x.unit_begin) * tx.n_obs * tx.n_variants;
            const uint64_t cy = (y.unit_end - y.unit_begin) * ty.n_obs * ty.n_variants;
            return cx > cy;
        });
        if (grp.items.size() > 0x7fffffffu) throw PlanError(PE_ERR_SIZE_LIMIT, "too many work items");
        if (!grp.items.empty()) grp.d_items.upload(grp.items);
    }
    L->partial_count = base;
    L->d_trades.upload(L->trades);
    L->d_obs.upload(obs);
    L->d_obs32.upload(obs32);
    if (!sobol.empty()) L->d_sobol.upload(sobol);
    if (!graph_ops.empty()) {
        L->d_graph_ops.upload(graph_ops);
        L->d_graph_segs.upload(graph_segs);
        L->d_graph_consts.upload(graph_consts);
    }
    L->partials.reserve(std::max<size_t>(1, base) * sizeof(BlockStat));
    L->err.reserve(sizeof(unsigned long long));
    L->host_partials.resize(base);
    p0.bs_launch.reset(L.release());
    return *p0.bs_launch;
}

// Enqueues the launch of every non-empty group in `mask` with the given
// persistent grid sizes (0: one workgroup per item).
BsBatchArgs bs_args(BsLaunch& L, int g, uint32_t grid) {
    BsGroup& grp = L.groups[g];
    BsBatchArgs a{};
    a.trades = L.d_trades.as<BsTrade>();
    a.obs = L.d_obs.as<BsObs>();
    a.obs32 = L.d_obs32.as<BsObs32>();
    a.sobol = L.d_sobol.as<uint32_t>();
    a.items = grp.d_items.as<BsItem>();
    a.partials = static_cast<BlockStat*>(L.partials.ptr) + grp.partial_base;
    a.error = static_cast<unsigned long long*>(L.err.ptr);
    a.graph_ops = L.d_graph_ops.as<GraphOp>();
    a.graph_segs = L.d_graph_segs.as<uint32_t>();
    a.graph_consts = L.d_graph_consts.as<double>();
    a.n_items = static_cast<uint32_t>(grp.items.size());
    a.components = grp.components;
    a.grid = grid;
    a.graph_regs = grp.graph_regs;
    return a;
}

uint32_t enqueue_bs_batch(BsLaunch& L, pe_backend backend, const uint32_t (&grid)[kBsGroups],
                          uint32_t mask = ~0u) {
    uint32_t launches = 0;
    for (int g = 0; g < kBsGroups; ++g) {
        BsGroup& grp = L.groups[g];
        if (grp.items.empty() || !((mask >> g) & 1u)) continue;
        const BsBatchArgs a = bs_args(L, g, grid[g]);
        const int kind = bs_group_kind(g);
        check(backend == PE_BACKEND_GPU_FP64
                  ? launch_bs_batch_f64(a, grp.risk, grp.cont, kind, grp.graph_regs, t_stream)
                  : launch_bs_batch_f32(a, grp.risk, grp.cont, kind, grp.graph_regs, t_stream),
              "bs kernel launch");
        ++launches;
    }
    return launches;
}

// Launch-shape tuner: the persistent grid size of each group is picked among
// one workgroup per item and 4, 8 or 16 resident workgroups per CU by timing
// the group on its own items (results are bit-identical across the
// candidates).  Batches of fewer than kTuneMinItems items keep one workgroup
// per item; PE_BATCH_TUNE=0 turns the tuner off.  The tuned launches are then
// captured into one HIP graph that later evaluations replay.
constexpr size_t kTuneMinItems = 4096;

void tune_bs_batch(BsLaunch& L, pe_backend backend) {
    L.tuned = true;
    const char* env = std::getenv("PE_BATCH_TUNE");
    if (env && std::strcmp(env, "0") == 0) return;
    int device = 0, cus = 0;
    check(hipGetDevice(&device), "hipGetDevice");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
    for (int g = 0; g < kBsGroups; ++g) {
        const size_t n = L.groups[g].items.size();
        if (n < kTuneMinItems) continue;
        float best = 0.0f;
        uint32_t best_grid = 0;
        for (uint32_t per_cu : {0u, 4u, 8u, 16u}) {
            const uint32_t candidate = per_cu * static_cast<uint32_t>(cus);
            if (per_cu && candidate >= n) continue;
            uint32_t grid[kBsGroups] = {};
            grid[g] = candidate;
            const float ms = timed([&] { enqueue_bs_batch(L, backend, grid, 1u << g); });
            if (per_cu == 0) best = ms;
            if (ms < best) {
                best = ms;
                best_grid = candidate;
            }
        }
        L.grid[g] = best_grid;
    }
}

void evaluate_bs_batch(std::vector<HostPlan*>& plans, pe_backend backend, pe_result* outs) {
    BsLaunch& L = prepare_bs(plans, backend);
    const unsigned long long none = ~0ull;
    if (!L.tuned) {
        check(hipMemcpyAsync(L.err.ptr, &none, sizeof(none), hipMemcpyHostToDevice, t_stream), "error word");
        tune_bs_batch(L, backend);
        // the tuned launches as kernel nodes of one graph, built explicitly
        // (no stream capture, so other threads' calls never meet a capture)
        check(hipGraphCreate(&L.graph, 0), "graph create");
        for (int g = 0; g < kBsGroups; ++g) {
            if (L.groups[g].items.empty()) continue;
            const BsBatchArgs a = bs_args(L, g, L.grid[g]);
            const BsGroup& grp = L.groups[g];
            const int kind = bs_group_kind(g);
            check(backend == PE_BACKEND_GPU_FP64
                      ? add_bs_batch_node_f64(L.graph, a, grp.risk, grp.cont, kind, grp.graph_regs)
                      : add_bs_batch_node_f32(L.graph, a, grp.risk, grp.cont, kind, grp.graph_regs),
                  "graph kernel node");
        }
        check(hipGraphInstantiate(&L.graph_exec, L.graph, nullptr, nullptr, 0), "graph instantiate");
    }
    check(hipMemcpyAsync(L.err.ptr, &none, sizeof(none), hipMemcpyHostToDevice, t_stream), "error word");
    uint32_t launches = 0;
    for (int g = 0; g < kBsGroups; ++g) launches += L.groups[g].items.empty() ? 0u : 1u;
    const float ms = timed([&] { check(hipGraphLaunch(L.graph_exec, t_stream), "graph launch"); });
    check(copy_on_stream(L.host_partials.data(), L.partials.ptr, L.partial_count * sizeof(BlockStat),
                    hipMemcpyDeviceToHost),
          "partials copy");
    unsigned long long err = none;
    check(copy_on_stream(&err, L.err.ptr, sizeof(err), hipMemcpyDeviceToHost), "error copy");
    if (err != none && (err >> 63)) {
        const uint64_t trade = (err >> 40) & ((1ull << 23) - 1), unit = err & ((1ull << 40) - 1);
        throw PlanError(PE_ERR_NONFINITE_STATE, "trade " + std::to_string(trade) + ", sampling unit " +
                                                    std::to_string(unit) + ": non-finite graph payoff value");
    }
    if (err != none) {
        const uint64_t trade = err >> 40, unit = err & ((1ull << 40) - 1);
        throw PlanError(PE_ERR_NONFINITE_STATE, "trade " + std::to_string(trade) + ", sampling unit " +
                                                    std::to_string(unit) +
                                                    ": invalid barrier bridge state (non-positive transformed barrier or state)");
    }
    for (size_t ti = 0; ti < plans.size(); ++ti) {
        const BsPlanSlot& slot = L.slots[ti];
        const BsGroup& grp = L.groups[slot.group];
        const pe_engine& e = plans[ti]->req.engine;
        const uint32_t scrambles = e.kind == PE_ENGINE_RQMC ? e.scrambles : 1u;
        const size_t first = grp.partial_base + static_cast<size_t>(slot.first_partial) * grp.components;
        std::vector<BlockStat> mine(L.host_partials.begin() + first,
                                    L.host_partials.begin() + first +
                                        static_cast<size_t>(scrambles) * slot.chunks * grp.components);
        pe_result* out = outs + ti;
        std::memset(out, 0, sizeof(*out));
        assemble(*plans[ti], mine, slot.chunks, scrambles, grp.components, 0, grp.components, out);
        out->device_ms = ms;
    }
    g_kernel_ms = ms;
    g_kernel_launches = launches;
}

// ------------------------------------------------- allocation-free path --
// One BS/Black-76 plan through the batch kernel with every device buffer --
// trade record, observation table, Sobol table, graph tape, work items,
// block partials and the error word -- carved out of this device's fixed
// workspace (pe_noalloc.hip).  No device or pinned-host allocation, no event
// creation; the workspace is shared by every call on the device, so calls
// are serialised by g_noalloc_mutex and each call rewrites what it reads.
std::mutex g_noalloc_mutex;

void evaluate_bs_noalloc(HostPlan& p, pe_backend backend, pe_result* out) {
    if (p.lv || p.rough || p.american)
        throw PlanError(PE_ERR_UNSUPPORTED, "the allocation-free entry point prices Black-Scholes / Black-76 plans");
    unsigned char* base = static_cast<unsigned char*>(scratch_pointer(g_device));
    if (!base) throw PlanError(PE_ERR_DEVICE, "workspace address not set up (pe_init)");
    const pe_engine& e = p.req.engine;
    const bool rqmc = e.kind == PE_ENGINE_RQMC;
    const uint32_t scrambles = rqmc ? e.scrambles : 1u;
    const int g = bs_group_of(p);
    const bool risk = bs_group_risk(g), cont = bs_group_cont(g);
    const uint32_t comps = cont ? PE_COMPONENTS : (risk ? 10u : 1u);
    const uint64_t chunks64 = (e.units + kItemUnits - 1) / kItemUnits;
    const uint64_t n_items64 = chunks64 * scrambles;
    if (rqmc && p.host_sobol.empty()) p.host_sobol = build_sobol_table(e.seed, e.scrambles, p.sobol_dims);
    // layout: error word | trade | observations | Sobol | graph ops | segs | consts | items | partials
    auto align = [](size_t x) { return (x + 255u) & ~static_cast<size_t>(255u); };
    size_t off = 0;
    const size_t o_err = off; off = align(off + sizeof(unsigned long long));
    const size_t o_trade = off; off = align(off + sizeof(BsTrade));
    const size_t o_obs = off; off = align(off + p.bs_obs.size() * sizeof(BsObs));
    const size_t o_obs32 = off; off = align(off + p.bs_obs32.size() * sizeof(BsObs32));
    const size_t o_sobol = off; off = align(off + p.host_sobol.size() * sizeof(uint32_t));
    const GraphTape& gt = p.graph_tape;
    const size_t o_gops = off; off = align(off + (p.graph ? gt.ops.size() * sizeof(GraphOp) : 0));
    const size_t o_gsegs = off; off = align(off + (p.graph ? gt.segs.size() * sizeof(uint32_t) : 0));
    const size_t o_gconsts = off; off = align(off + (p.graph ? gt.consts.size() * sizeof(double) : 0));
    const size_t o_items = off;
    const size_t need = o_items + align(n_items64 * sizeof(BsItem)) + n_items64 * comps * sizeof(BlockStat);
    if (n_items64 > 0x7fffffffull || need > scratch_bytes())
        throw PlanError(PE_ERR_SIZE_LIMIT, "the plan needs " + std::to_string(need) +
                                               " bytes of device workspace; the allocation-free entry point has " +
                                               std::to_string(scratch_bytes()));
    const uint32_t chunks = static_cast<uint32_t>(chunks64), n_items = static_cast<uint32_t>(n_items64);
    off = align(o_items + n_items64 * sizeof(BsItem));
    const size_t o_partials = off;
    BsTrade t = p.trade;
    t.obs_offset = 0;
    t.sobol_offset = 0;
    if (p.graph) {
        t.graph_ops = t.graph_segs = t.graph_consts = 0;
        if (gt.segs.size() != p.bs_obs.size() + 2) throw PlanError(PE_ERR_INVALID_ARGUMENT, "graph schedule");
    }
    // items in unit order (one workgroup per item, partial = item index)
    std::vector<BsItem> items(n_items);
    for (uint32_t sc = 0; sc < scrambles; ++sc)
        for (uint32_t c = 0; c < chunks; ++c) {
            BsItem& it = items[static_cast<size_t>(sc) * chunks + c];
            it = BsItem{};
            it.scramble = sc;
            it.unit_begin = static_cast<uint64_t>(c) * kItemUnits;
            it.unit_end = std::min<uint64_t>(it.unit_begin + kItemUnits, e.units);
            it.partial = sc * chunks + c;
        }
    auto put = [&](size_t at, const void* src, size_t bytes) {
        if (bytes) check(hipMemcpyAsync(base + at, src, bytes, hipMemcpyHostToDevice, t_stream), "workspace upload");
    };
    const unsigned long long none = ~0ull;
    put(o_err, &none, sizeof(none));
    put(o_trade, &t, sizeof(t));
    put(o_obs, p.bs_obs.data(), p.bs_obs.size() * sizeof(BsObs));
    put(o_obs32, p.bs_obs32.data(), p.bs_obs32.size() * sizeof(BsObs32));
    put(o_sobol, p.host_sobol.data(), p.host_sobol.size() * sizeof(uint32_t));
    if (p.graph) {
        put(o_gops, gt.ops.data(), gt.ops.size() * sizeof(GraphOp));
        put(o_gsegs, gt.segs.data(), gt.segs.size() * sizeof(uint32_t));
        put(o_gconsts, gt.consts.data(), gt.consts.size() * sizeof(double));
    }
    put(o_items, items.data(), items.size() * sizeof(BsItem));
    BsBatchArgs a{};
    a.trades = reinterpret_cast<const BsTrade*>(base + o_trade);
    a.obs = reinterpret_cast<const BsObs*>(base + o_obs);
    a.obs32 = reinterpret_cast<const BsObs32*>(base + o_obs32);
    a.sobol = reinterpret_cast<const uint32_t*>(base + o_sobol);
    a.items = reinterpret_cast<const BsItem*>(base + o_items);
    a.partials = base + o_partials;
    a.error = reinterpret_cast<unsigned long long*>(base + o_err);
    a.graph_ops = reinterpret_cast<const GraphOp*>(base + o_gops);
    a.graph_segs = reinterpret_cast<const uint32_t*>(base + o_gsegs);
    a.graph_consts = reinterpret_cast<const double*>(base + o_gconsts);
    a.n_items = n_items;
    a.components = comps;
    a.grid = 0;
    const uint32_t graph_regs = p.graph ? gt.regs : 0u;
    a.graph_regs = graph_regs;
    const int kind = bs_group_kind(g);
    check(backend == PE_BACKEND_GPU_FP64 ? launch_bs_batch_f64(a, risk, cont, kind, graph_regs, t_stream)
                                         : launch_bs_batch_f32(a, risk, cont, kind, graph_regs, t_stream),
          "bs kernel launch");
    std::vector<BlockStat> partials(static_cast<size_t>(n_items) * comps);
    unsigned long long err = none;
    check(hipMemcpyAsync(partials.data(), base + o_partials, partials.size() * sizeof(BlockStat),
                         hipMemcpyDeviceToHost, t_stream),
          "partials copy");
    check(hipMemcpyAsync(&err, base + o_err, sizeof(err), hipMemcpyDeviceToHost, t_stream), "error copy");
    check(hipStreamSynchronize(t_stream), "stream synchronize");
    if (err != none) {
        const uint64_t unit = err & ((1ull << 40) - 1);
        out->failing_unit = unit;
        throw PlanError(PE_ERR_NONFINITE_STATE,
                        "sampling unit " + std::to_string(unit) +
                            ((err >> 63) ? ": non-finite graph payoff value"
                                         : ": invalid barrier bridge state (non-positive transformed barrier or state)"));
    }
    std::memset(out, 0, sizeof(*out));
    assemble(p, partials, chunks, scrambles, comps, 0, comps, out);
    g_kernel_ms = 0.0;
    g_kernel_launches = 1;
}

// ---------------------------------------------------------- lifted Heston --
// One simulated path set for every plan: the union of the plans' events is
// the grid (one plan: the reference's own grid).  The path kernel stores the
// physical spot at each payoff's expiry node; the payoff kernel reduces.
}  // namespace

struct RoughLaunch {
    std::vector<uint64_t> key;
    int backend = 0;
    RoughArgs a{};
    RoughPayoffArgs pa{};
    dim3 grid;
    uint32_t n = 0, chunks = 0, scrambles = 1;
    size_t items = 0;
    bool rqmc = false;
    DeviceBuffer d_steps, d_den, d_rden, d_store_nodes, d_store, d_sobol, d_err, d_products, d_items, d_partials;
    std::vector<BlockStat> host;
};

void RoughLaunchDeleter::operator()(RoughLaunch* l) const { delete l; }

// Device state of a Bermudan under lifted Heston (one plan per launch).
struct RoughLsmLaunch {
    int backend = 0;
    RoughLsmArgs args{};
    dim3 train_grid, value_grid;
    uint32_t n = 0, chunks = 0, scrambles = 1, dates = 0;
    DeviceBuffer d_steps, d_den, d_rden, d_train_sobol, d_value_sobol, d_nodes, d_discount, d_scale, store_s, store_v,
        cash, partials, scalings, decisions, value_partials, counts, stops, in_sample, err;
    std::vector<BlockStat> host_partials;
    std::vector<RoughLsmDecision> host_decisions;
    std::vector<unsigned long long> host_counts;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // start, paths done, training done, valuation done
    // network policy (PE_LSM_MLP): the polynomial policy above is its judge
    bool nn = false;
    RoughNnArgs nn_args{};
    DeviceBuffer d_times, nn_cash, nn_init, nn_weights, nn_adam_m, nn_adam_v, nn_rows, nn_grad, nn_stats, nn_offsets,
        nn_norms, nn_in_sample, nn_partials, nn_counts;
    std::vector<BlockStat> host_nn_partials;
    std::vector<unsigned long long> host_nn_counts;
    std::vector<RoughNnNorm> host_norms;
    unsigned long long rows_trained = 0;
    hipEvent_t nn_ev[3] = {nullptr, nullptr, nullptr};   // network start, trained, valued
    ~RoughLsmLaunch() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : nn_ev)
            if (e) (void)hipEventDestroy(e);
    }
};

void RoughLsmLaunchDeleter::operator()(RoughLsmLaunch* l) const { delete l; }

namespace {

void prepare_rough(std::vector<HostPlan*>& plans, pe_backend backend);

void evaluate_rough(std::vector<HostPlan*>& plans, pe_backend backend, pe_result* outs) {
    HostPlan& p0 = *plans[0];
    std::vector<uint64_t> key;
    for (const HostPlan* q : plans) key.push_back(q->id);
    if (!(p0.rough_launch && p0.rough_launch->key == key && p0.rough_launch->backend == static_cast<int>(backend)))
        prepare_rough(plans, backend);
    RoughLaunch& L = *p0.rough_launch;
    const unsigned long long none = ~0ull;
    check(hipMemcpyAsync(L.d_err.ptr, &none, sizeof(none), hipMemcpyHostToDevice, t_stream), "error word");
    const float ms = timed([&] {
        check(backend == PE_BACKEND_GPU_FP64
                  ? launch_rough_f64(L.a, L.pa, L.grid, static_cast<uint32_t>(L.items),
                                     static_cast<unsigned long long*>(L.d_err.ptr), t_stream)
                  : launch_rough_f32(L.a, L.pa, L.grid, static_cast<uint32_t>(L.items),
                                     static_cast<unsigned long long*>(L.d_err.ptr), t_stream),
              "rough kernel launch");
    });
    check(copy_on_stream(L.host.data(), L.d_partials.ptr, L.host.size() * sizeof(BlockStat), hipMemcpyDeviceToHost),
          "partials");
    unsigned long long err = none;
    check(copy_on_stream(&err, L.d_err.ptr, sizeof(err), hipMemcpyDeviceToHost), "error word");
    if (err != none)
        throw PlanError(PE_ERR_NONFINITE_STATE, "sampling unit " + std::to_string(err) + ": non-finite lifted Heston path");
    const size_t per = static_cast<size_t>(L.scrambles) * L.chunks;
    for (size_t k = 0; k < plans.size(); ++k) {
        std::vector<BlockStat> mine(L.host.begin() + static_cast<long>(k * per),
                                    L.host.begin() + static_cast<long>((k + 1) * per));
        pe_result* out = outs + k;
        std::memset(out, 0, sizeof(*out));
        assemble(*plans[k], mine, L.chunks, L.scrambles, 1, 0, 1, out);
        out->time_steps = L.n;
        out->path_steps = out->evaluated_paths * L.n;
        out->sobol_dimension = L.rqmc ? 2u * L.n : 0u;
        out->device_ms = ms;
    }
    g_kernel_ms = ms;
    g_kernel_launches = 2;
}

// Physical spot per unit of F at time t (HullWhiteDividendPlan node scale
// without dividends, crates/pricing/src/models/hull_white_dividends.rs:228-235).
double rough_scale(const HostPlan& p, double t) {
    const double pd = std::exp(p.discount_curve.log_discount(t));
    const double qd = std::exp(p.dividend_curve.log_discount(t));
    return 1.0 * p.spot / p.spot * qd / pd;
}

// The grid and model fields of RoughArgs: per-step dt, sqrt(dt) and the
// right node's spot scale, the [step][slot] tables 1 + x_i dt and their
// reciprocals, the lifted-Heston parameters.  Each step's row has the
// kernel's factor slots (rough_factor_slots); the slots past the factors
// hold +inf and 0, so those factors stay exactly zero and the kernels run
// every slot without per-factor branches.
void rough_grid_args(const HostPlan& p, const std::vector<double>& nodes, DeviceBuffer& d_steps, DeviceBuffer& d_den,
                     DeviceBuffer& d_rden, RoughArgs& a) {
    const HostLifted& h = p.lifted;
    const uint32_t n = static_cast<uint32_t>(nodes.size() - 1);
    const uint32_t nf = static_cast<uint32_t>(h.weights.size());
    const uint32_t slots = rough_factor_slots(nf);
    if (slots == 0) throw PlanError(PE_ERR_SIZE_LIMIT, "lifted Heston factor count");
    std::vector<RoughStep> steps(n);
    std::vector<double> den(static_cast<size_t>(n) * slots, std::numeric_limits<double>::infinity());
    std::vector<double> rden(den.size(), 0.0);
    for (uint32_t j = 0; j < n; ++j) {
        const double dt = nodes[j + 1] - nodes[j];
        steps[j] = RoughStep{dt, std::sqrt(dt), rough_scale(p, nodes[j + 1]), 0.0};
        for (uint32_t i = 0; i < nf; ++i) {
            const double d = 1.0 + h.rates[i] * dt;
            if (!std::isfinite(d)) throw PlanError(PE_ERR_INVALID_ARGUMENT, "lifted time scale overflow");
            den[static_cast<size_t>(j) * slots + i] = d;
            rden[static_cast<size_t>(j) * slots + i] = 1.0 / d;
        }
    }
    d_steps.upload(steps);
    d_den.upload(den);
    d_rden.upload(rden);
    a.steps = d_steps.as<RoughStep>();
    a.denominators = d_den.as<double>();
    a.reciprocals = d_rden.as<double>();
    a.spot = p.spot;
    a.scale0 = rough_scale(p, 0.0);
    a.v0 = h.v0;
    a.kappa = h.kappa;
    a.theta = h.theta;
    a.nu = h.nu;
    a.rho = h.rho;
    a.rho_bar = std::sqrt(std::max(1.0 - h.rho * h.rho, 0.0));
    for (uint32_t i = 0; i < nf; ++i) a.weights[i] = h.weights[i];
    a.n_steps = n;
    a.n_factors = nf;
}

void prepare_rough(std::vector<HostPlan*>& plans, pe_backend backend) {
    HostPlan& p0 = *plans[0];
    const pe_engine& e = p0.req.engine;
    std::vector<double> events;
    for (const HostPlan* q : plans) {
        if (!q->rough) throw PlanError(PE_ERR_INVALID_ARGUMENT, "mixing rough and non-rough plans");
        const pe_engine& f = q->req.engine;
        if (f.kind != e.kind || f.seed != e.seed || f.units != e.units || f.scrambles != e.scrambles ||
            f.antithetic != e.antithetic || f.brownian_bridge != e.brownian_bridge || q->spot != p0.spot ||
            q->rough_max_step != p0.rough_max_step || q->lifted.weights != p0.lifted.weights ||
            q->lifted.rates != p0.lifted.rates || q->lifted.v0 != p0.lifted.v0 ||
            q->lifted.kappa != p0.lifted.kappa || q->lifted.theta != p0.lifted.theta ||
            q->lifted.nu != p0.lifted.nu || q->lifted.rho != p0.lifted.rho)
            throw PlanError(PE_ERR_INVALID_ARGUMENT, "shared rough plans need one model, engine, spot and step");
        events.insert(events.end(), q->rough_events.begin(), q->rough_events.end());
    }
    if (e.brownian_bridge) throw PlanError(PE_ERR_UNSUPPORTED, "Brownian-bridge ordering of lifted Heston paths on the device");
    const std::vector<double> nodes = compile_time_grid(events, p0.rough_max_step);
    const uint32_t n = static_cast<uint32_t>(nodes.size() - 1);
    if (n > 65536) throw PlanError(PE_ERR_SIZE_LIMIT, "rough time grid resource limit");
    // store slots: the distinct expiry nodes of the plans
    std::vector<uint32_t> store_nodes;
    std::vector<uint32_t> plan_node(plans.size());
    for (size_t k = 0; k < plans.size(); ++k) {
        const auto it = std::lower_bound(nodes.begin(), nodes.end(), plans[k]->expiry);
        if (it == nodes.end() || *it != plans[k]->expiry) throw PlanError(PE_ERR_INVALID_ARGUMENT, "missing expiry node");
        plan_node[k] = static_cast<uint32_t>(it - nodes.begin());
        store_nodes.push_back(plan_node[k]);
    }
    std::sort(store_nodes.begin(), store_nodes.end());
    store_nodes.erase(std::unique(store_nodes.begin(), store_nodes.end()), store_nodes.end());
    const bool rqmc = e.kind == PE_ENGINE_RQMC;
    const uint32_t scrambles = rqmc ? e.scrambles : 1u;
    const int passes = e.antithetic ? 2 : 1;
    const uint64_t total_paths = e.units * scrambles * passes;
    const bool fp64 = backend == PE_BACKEND_GPU_FP64;
    std::unique_ptr<RoughLaunch, RoughLaunchDeleter> L(new RoughLaunch());
    L->key.clear();
    for (const HostPlan* q : plans) L->key.push_back(q->id);
    L->backend = static_cast<int>(backend);
    DeviceBuffer &d_steps = L->d_steps, &d_den = L->d_den, &d_rden = L->d_rden, &d_store_nodes = L->d_store_nodes,
                 &d_store = L->d_store, &d_sobol = L->d_sobol, &d_err = L->d_err, &d_products = L->d_products,
                 &d_items = L->d_items, &d_partials = L->d_partials;
    RoughArgs a{};
    rough_grid_args(p0, nodes, d_steps, d_den, d_rden, a);
    d_store_nodes.upload(store_nodes);
    d_store.reserve(total_paths * store_nodes.size() * (fp64 ? sizeof(double) : sizeof(float)));
    if (rqmc) d_sobol.upload(build_sobol_table(e.seed, e.scrambles, 2u * n));
    a.sobol = d_sobol.as<uint32_t>();
    a.store_nodes = d_store_nodes.as<uint32_t>();
    a.store = d_store.ptr;
    a.seed = e.seed;
    a.units = e.units;
    a.total_paths = total_paths;
    a.n_store = static_cast<uint32_t>(store_nodes.size());
    a.rqmc = rqmc ? 1u : 0u;
    a.antithetic = e.antithetic ? 1u : 0u;
    // payoff work: (product, scramble, 4096-unit chunk)
    std::vector<RoughProduct> products(plans.size());
    std::vector<RoughItem> items;
    const uint32_t chunks = static_cast<uint32_t>((e.units + 4095) / 4096);
    for (size_t k = 0; k < plans.size(); ++k) {
        const ProductDesc& d = plans[k]->product;
        RoughProduct& rp = products[k];
        rp.kind = d.kind;
        rp.call = d.call;
        rp.payout_asset = d.payout_asset;
        rp.slot = static_cast<uint32_t>(std::lower_bound(store_nodes.begin(), store_nodes.end(), plan_node[k]) -
                                        store_nodes.begin());
        rp.strike = d.strike;
        rp.notional = d.notional;
        rp.payout = d.payout;
        rp.discount = d.discount;
        for (uint32_t sc = 0; sc < scrambles; ++sc)
            for (uint32_t c = 0; c < chunks; ++c) {
                RoughItem it{};
                it.product = static_cast<uint32_t>(k);
                it.scramble = sc;
                it.unit_begin = static_cast<uint64_t>(c) * 4096;
                it.unit_end = std::min<uint64_t>(it.unit_begin + 4096, e.units);
                it.partial = static_cast<uint32_t>((k * scrambles + sc) * chunks + c);
                items.push_back(it);
            }
    }
    d_products.upload(products);
    d_items.upload(items);
    d_partials.reserve(items.size() * sizeof(BlockStat));
    RoughPayoffArgs pa{};
    pa.items = d_items.as<RoughItem>();
    pa.products = d_products.as<RoughProduct>();
    pa.store = d_store.ptr;
    pa.partials = d_partials.ptr;
    pa.units = e.units;
    pa.total_paths = total_paths;
    pa.antithetic = a.antithetic;
    d_err.reserve(sizeof(unsigned long long));
    const uint64_t blocks = (e.units + kBlockThreads - 1) / kBlockThreads;
    L->grid = dim3(static_cast<uint32_t>(std::min<uint64_t>(blocks, 1u << 14)), scrambles, 1u);
    L->a = a;
    L->pa = pa;
    L->n = n;
    L->chunks = chunks;
    L->scrambles = scrambles;
    L->items = items.size();
    L->rqmc = rqmc;
    L->host.resize(items.size());
    (void)fp64;
    p0.rough_launch.reset(L.release());
}

// The network policy of a Bermudan under lifted Heston (pe_rough_nn.inc):
// one network per non-terminal date, trained on the polynomial policy's FP64
// training stores with its own copy of the cash flows; 64 gradient
// workgroups of 128 rows per Adam step (batch a multiple of 128).
void prepare_rough_nn(const HostPlan& p, RoughLsmLaunch& L, uint32_t value_items) {
    const pe_lsm& l = p.req.lsm;
    if (l.mlp_hidden != 64) throw PlanError(PE_ERR_UNSUPPORTED, "MLP hidden width must be 64 on the device");
    if (l.mlp_epochs == 0 || !(l.mlp_learning_rate > 0.0)) throw PlanError(PE_ERR_INVALID_ARGUMENT, "MLP epochs / rate");
    const uint32_t B = l.mlp_batch;
    if (B == 0 || B % 128 != 0 || B > (1u << 20))
        throw PlanError(PE_ERR_INVALID_ARGUMENT, "MLP batch must be a positive multiple of 128, at most 2^20");
    const uint32_t G = std::max<uint32_t>(1, std::min<uint32_t>(64, B / 128));
    if (B % (128 * G) != 0) throw PlanError(PE_ERR_INVALID_ARGUMENT, "MLP batch must split into 128-row tiles");
    const RoughLsmArgs& a = L.args;
    const uint32_t dates = a.n_dates;
    const size_t stride = (kMlpParameterCount + 3) / 4 * 4;
    L.nn = true;
    std::vector<double> times(p.obs_times.begin(), p.obs_times.end());
    L.d_times.upload(times);
    std::vector<float> init(kMlpParameterCount, 0.0f);
    uint64_t state = l.mlp_seed;
    auto uniform = [&]() {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return 2.0 * (static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    };
    for (int k = 0; k < 128; ++k) init[k] = static_cast<float>(uniform() * std::sqrt(6.0 / 2.0));
    for (int k = 192; k < 4288; ++k) init[k] = static_cast<float>(uniform() * std::sqrt(6.0 / 64.0));
    L.nn_init.upload(init);   // output layer at zero: the untrained network is the control variate
    const size_t nets = std::max<uint32_t>(1, dates - 1);
    L.nn_cash.reserve(a.train_paths * sizeof(double));
    L.nn_weights.reserve(nets * stride * sizeof(float));
    L.nn_adam_m.reserve(stride * sizeof(float));
    L.nn_adam_v.reserve(stride * sizeof(float));
    L.nn_rows.reserve(std::max<uint64_t>(1, a.train_paths) * 4 * sizeof(float));
    L.nn_grad.reserve(static_cast<size_t>(G) * (kMlpParameterCount + 1) * sizeof(float));
    L.nn_stats.reserve(static_cast<size_t>(a.n_wg) * 8 * sizeof(double));
    L.nn_offsets.reserve(static_cast<size_t>(a.n_wg) * sizeof(uint32_t));
    L.nn_norms.reserve(nets * sizeof(RoughNnNorm));
    L.nn_in_sample.reserve(sizeof(double));
    L.nn_partials.reserve(static_cast<size_t>(value_items) * sizeof(BlockStat));
    L.nn_counts.reserve(dates * sizeof(unsigned long long));
    L.host_nn_partials.resize(value_items);
    L.host_nn_counts.resize(dates);
    L.host_norms.resize(nets);
    RoughNnArgs& A = L.nn_args;
    A.ex_nodes = a.ex_nodes;
    A.ex_discount = a.ex_discount;
    A.ex_scale = a.ex_scale;
    A.ex_times = L.d_times.as<double>();
    A.store_s = a.store_s;
    A.store_v = a.store_v;
    A.cash = L.nn_cash.as<double>();
    A.init = L.nn_init.as<float>();
    A.weights = L.nn_weights.as<float>();
    A.adam_m = L.nn_adam_m.as<float>();
    A.adam_v = L.nn_adam_v.as<float>();
    A.rows = L.nn_rows.as<float4>();
    A.grad = L.nn_grad.as<float>();
    A.stats = L.nn_stats.as<double>();
    A.offsets = L.nn_offsets.as<uint32_t>();
    A.norms = L.nn_norms.as<RoughNnNorm>();
    A.in_sample = L.nn_in_sample.as<double>();
    A.value_partials = L.nn_partials.ptr;
    A.counts = L.nn_counts.as<unsigned long long>();
    A.error = a.error;
    A.train_paths = a.train_paths;
    A.value_chunk = a.value_chunk;
    A.strike = a.strike;
    A.notional = a.notional;
    A.itm_tol = a.itm_tol;
    A.kappa = p.lifted.kappa;
    A.theta = p.lifted.theta;
    A.lr = static_cast<float>(l.mlp_learning_rate);
    A.beta1 = 0.9f;
    A.beta2 = 0.999f;
    A.eps = 1e-8f;
    A.n_dates = dates;
    A.call = a.call;
    A.n_wg = a.n_wg;
    A.batch = B;
    A.epochs = l.mlp_epochs;
    A.grad_wgs = G;
    A.rows_per_wg = B / G;
    A.value_items = value_items;
    for (hipEvent_t& e : L.nn_ev) check(hipEventCreate(&e), "hipEventCreate");
}

// A Bermudan under lifted Heston: the training paths, the backward
// induction over the exercise dates and the valuation (pe_rough_lsm.inc).
void prepare_rough_lsm(HostPlan& p, pe_backend backend) {
    if (p.rough_lsm_launch && p.rough_lsm_launch->backend == static_cast<int>(backend)) return;
    std::unique_ptr<RoughLsmLaunch, RoughLsmLaunchDeleter> L(new RoughLsmLaunch());
    L->backend = static_cast<int>(backend);
    const pe_engine& ve = p.req.engine;
    const pe_engine& te = p.req.lsm.training_engine;
    const std::vector<double> nodes = compile_time_grid(p.rough_events, p.rough_max_step);
    const uint32_t n = static_cast<uint32_t>(nodes.size() - 1);
    const uint32_t dates = static_cast<uint32_t>(p.obs_times.size());
    RoughLsmArgs& a = L->args;
    RoughArgs g{};
    rough_grid_args(p, nodes, L->d_steps, L->d_den, L->d_rden, g);
    std::vector<uint32_t> ex_nodes(dates);
    std::vector<double> ex_df(dates), ex_scale(dates);
    for (uint32_t k = 0; k < dates; ++k) {
        const auto it = std::lower_bound(nodes.begin(), nodes.end(), p.obs_times[k]);
        if (it == nodes.end() || *it != p.obs_times[k] || it == nodes.begin())
            throw PlanError(PE_ERR_INVALID_ARGUMENT, "exercise dates must be future grid nodes");
        ex_nodes[k] = static_cast<uint32_t>(it - nodes.begin());
        ex_df[k] = std::exp(p.discount_curve.log_discount(p.obs_times[k]));
        ex_scale[k] = rough_scale(p, p.obs_times[k]);
    }
    L->d_nodes.upload(ex_nodes);
    L->d_discount.upload(ex_df);
    L->d_scale.upload(ex_scale);
    auto engine = [&](const pe_engine& e, uint32_t domain, DeviceBuffer& sobol) {
        RoughArgs r = g;
        r.seed = e.seed;
        r.units = e.units;
        r.rqmc = e.kind == PE_ENGINE_RQMC ? 1u : 0u;
        r.antithetic = e.antithetic ? 1u : 0u;
        r.domain = domain;
        if (r.rqmc) {
            sobol.upload(build_sobol_table(e.seed, e.scrambles, 2u * n));
            r.sobol = sobol.as<uint32_t>();
        }
        return r;
    };
    a.train = engine(te, 1u, L->d_train_sobol);   // RandomDomain::LsmTrain
    a.value = engine(ve, 0u, L->d_value_sobol);   // RandomDomain::Valuation
    const uint32_t train_scr = te.kind == PE_ENGINE_RQMC ? te.scrambles : 1u;
    const uint32_t value_scr = ve.kind == PE_ENGINE_RQMC ? ve.scrambles : 1u;
    a.train_paths = te.units * train_scr * (te.antithetic ? 2u : 1u);
    a.ex_nodes = L->d_nodes.as<uint32_t>();
    a.ex_discount = L->d_discount.as<double>();
    a.ex_scale = L->d_scale.as<double>();
    L->store_s.reserve(static_cast<size_t>(dates) * a.train_paths * sizeof(double));
    L->store_v.reserve(static_cast<size_t>(dates) * a.train_paths * sizeof(double));
    L->cash.reserve(a.train_paths * sizeof(double));
    a.store_s = L->store_s.as<double>();
    a.store_v = L->store_v.as<double>();
    a.cash = L->cash.as<double>();
    // training passes: about 512 paths per workgroup (two per lane), at most
    // 4096 workgroups
    a.n_wg = static_cast<uint32_t>(std::min<uint64_t>(4096, std::max<uint64_t>(1, (a.train_paths + 511) / 512)));
    L->partials.reserve(static_cast<size_t>(a.n_wg) * kRlGram * sizeof(double));
    a.partials = L->partials.as<double>();
    L->scalings.reserve(std::max<size_t>(1, dates) * sizeof(RoughLsmScaling));
    L->decisions.reserve(std::max<size_t>(1, dates) * sizeof(RoughLsmDecision));
    a.scalings = L->scalings.as<RoughLsmScaling>();
    a.decisions = L->decisions.as<RoughLsmDecision>();
    // valuation: one workgroup per (chunk of units, scramble), a multiple of
    // 256 units per chunk and about 2048 workgroups in all
    const uint64_t value_units = ve.units * value_scr;
    a.value_chunk = 256u * std::max<uint64_t>(1, (value_units + 256u * 2048u - 1) / (256u * 2048u));
    const uint32_t chunks = static_cast<uint32_t>((ve.units + a.value_chunk - 1) / a.value_chunk);
    L->value_partials.reserve(static_cast<size_t>(chunks) * value_scr * sizeof(BlockStat));
    a.value_partials = L->value_partials.ptr;
    L->counts.reserve(dates * sizeof(unsigned long long));
    a.counts = L->counts.as<unsigned long long>();
    L->in_sample.reserve(sizeof(double));
    a.in_sample = L->in_sample.as<double>();
    L->err.reserve(sizeof(unsigned long long));
    a.error = L->err.as<unsigned long long>();
    a.strike = p.req.product.strike;
    a.notional = p.req.product.notional;
    a.itm_tol = p.req.lsm.itm_abs_tolerance;
    a.abs_tol = p.req.lsm.abs_rank_tolerance;
    a.rel_tol = p.req.lsm.rel_rank_tolerance;
    a.n_dates = dates;
    a.call = p.req.product.side == PE_CALL ? 1u : 0u;
    a.columns = p.req.lsm.max_degree == 1 ? 3u : 6u;
    a.value_items = chunks * value_scr;
    const uint64_t train_blocks = (te.units + kBlockThreads - 1) / kBlockThreads;
    L->train_grid = dim3(static_cast<uint32_t>(std::min<uint64_t>(train_blocks, 1u << 14)), train_scr, 1u);
    L->value_grid = dim3(chunks, value_scr, 1u);
    L->n = n;
    L->chunks = chunks;
    L->scrambles = value_scr;
    L->dates = dates;
    L->host_partials.resize(static_cast<size_t>(chunks) * value_scr);
    L->host_decisions.resize(dates);
    L->host_counts.resize(dates);
    for (hipEvent_t& e : L->ev) check(hipEventCreate(&e), "hipEventCreate");
    if (p.req.lsm.method == PE_LSM_MLP) prepare_rough_nn(p, *L, chunks * value_scr);
    p.rough_lsm_launch = std::move(L);
}

void evaluate_rough_lsm(HostPlan& p, pe_backend backend, pe_result* out, pe_lsm_report* report) {
    prepare_rough_lsm(p, backend);
    RoughLsmLaunch& L = *p.rough_lsm_launch;
    RoughLsmArgs a = L.args;
    const uint64_t vtraj = a.value.units * L.scrambles * (a.value.antithetic ? 2u : 1u);
    const bool want_stops = report && report->stopping_indices && report->stopping_capacity;
    if (want_stops) {
        L.stops.reserve(std::max<uint64_t>(1, vtraj) * sizeof(uint32_t));
        a.stops = L.stops.as<uint32_t>();
    }
    const unsigned long long none = ~0ull;
    check(hipMemcpyAsync(L.err.ptr, &none, sizeof(none), hipMemcpyHostToDevice, t_stream), "error word");
    const std::vector<unsigned long long> zeros(L.dates, 0ull);
    check(hipMemcpyAsync(L.counts.ptr, zeros.data(), L.dates * sizeof(unsigned long long), hipMemcpyHostToDevice,
                         t_stream),
          "counts");
    uint32_t* nn_stops = nullptr;   // the network policy's stopping indices are the ones reported
    if (L.nn) {
        nn_stops = a.stops;
        a.stops = nullptr;
    }
    check(hipEventRecord(L.ev[0], t_stream), "hipEventRecord");
    void* nn_cash = L.nn ? L.nn_cash.ptr : nullptr;
    check(backend == PE_BACKEND_GPU_FP64
              ? launch_rough_lsm_f64(a, L.train_grid, L.value_grid, t_stream, L.ev[1], L.ev[2], nn_cash)
              : launch_rough_lsm_f32(a, L.train_grid, L.value_grid, t_stream, L.ev[1], L.ev[2], nn_cash),
          "rough LSM launch");
    check(hipEventRecord(L.ev[3], t_stream), "hipEventRecord");
    if (L.nn) {
        // the network policy: per date a sweep + reduction, the row count
        // read back, then compaction and the Adam steps; then date 0's
        // in-sample pass and the valuation
        RoughNnArgs A = L.nn_args;
        A.stops = nn_stops;
        check(hipMemsetAsync(L.nn_counts.ptr, 0, L.dates * sizeof(unsigned long long), t_stream), "counts");
        check(hipEventRecord(L.nn_ev[0], t_stream), "hipEventRecord");
        uint32_t adam_t = 0;
        L.rows_trained = 0;
        check(hipMemsetAsync(L.nn_adam_m.ptr, 0, L.nn_adam_m.bytes, t_stream), "Adam state");
        check(hipMemsetAsync(L.nn_adam_v.ptr, 0, L.nn_adam_v.bytes, t_stream), "Adam state");
        for (int d = static_cast<int>(L.dates) - 2; d >= 0; --d) {
            const uint32_t dd = static_cast<uint32_t>(d);
            check(launch_rough_nn_date(A, dd, t_stream), "network sweep");
            RoughNnNorm nm{};
            check(copy_on_stream(&nm, L.nn_norms.as<RoughNnNorm>() + dd, sizeof(nm), hipMemcpyDeviceToHost),
                  "row count copy");
            if (nm.n == 0) continue;
            const uint32_t steps = A.epochs * ((nm.n + A.batch - 1) / A.batch);
            check(launch_rough_nn_train(A, dd, steps, adam_t, t_stream), "network training");
            adam_t += steps;
            L.rows_trained += static_cast<unsigned long long>(A.epochs) * nm.n;
        }
        check(launch_rough_nn_final(A, t_stream), "network in-sample pass");
        check(hipEventRecord(L.nn_ev[1], t_stream), "hipEventRecord");
        check(backend == PE_BACKEND_GPU_FP64 ? launch_rough_nn_value_f64(A, a.value, L.value_grid, t_stream)
                                             : launch_rough_nn_value_f32(A, a.value, L.value_grid, t_stream),
              "network valuation");
        check(hipEventRecord(L.nn_ev[2], t_stream), "hipEventRecord");
        check(hipEventSynchronize(L.nn_ev[2]), "hipEventSynchronize");
    } else {
        check(hipEventSynchronize(L.ev[3]), "hipEventSynchronize");
    }
    float path_ms = 0.0f, train_ms = 0.0f, value_ms = 0.0f, nn_train_ms = 0.0f, nn_value_ms = 0.0f;
    check(hipEventElapsedTime(&path_ms, L.ev[0], L.ev[1]), "hipEventElapsedTime");
    check(hipEventElapsedTime(&train_ms, L.ev[0], L.ev[2]), "hipEventElapsedTime");
    check(hipEventElapsedTime(&value_ms, L.ev[2], L.ev[3]), "hipEventElapsedTime");
    if (L.nn) {
        check(hipEventElapsedTime(&nn_train_ms, L.nn_ev[0], L.nn_ev[1]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&nn_value_ms, L.nn_ev[1], L.nn_ev[2]), "hipEventElapsedTime");
    }
    unsigned long long err = none;
    check(copy_on_stream(&err, L.err.ptr, sizeof(err), hipMemcpyDeviceToHost), "error copy");
    if (err != none) {
        out->failing_unit = err;
        throw PlanError(PE_ERR_NONFINITE_STATE, "sampling unit " + std::to_string(err) + ": non-finite lifted Heston path");
    }
    check(copy_on_stream(L.host_partials.data(), L.value_partials.ptr, L.host_partials.size() * sizeof(BlockStat),
                         hipMemcpyDeviceToHost),
          "partials copy");
    check(copy_on_stream(L.host_counts.data(), L.counts.ptr, L.dates * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost),
          "counts copy");
    if (L.dates > 1)
        check(copy_on_stream(L.host_decisions.data(), L.decisions.ptr, (L.dates - 1) * sizeof(RoughLsmDecision),
                             hipMemcpyDeviceToHost),
              "decisions copy");
    double in_sample = 0.0;
    check(copy_on_stream(&in_sample, L.in_sample.ptr, sizeof(double), hipMemcpyDeviceToHost), "in-sample copy");
    std::memset(out, 0, sizeof(*out));
    assemble(p, L.host_partials, L.chunks, L.scrambles, 1, 0, 1, out);
    // network policy: its out-of-sample value against the polynomial judge's
    // on the same valuation paths; the higher one is the price
    pe_result judge = *out;
    bool fallback = false;
    double nn_in_sample = 0.0;
    pe_estimate network_est{};
    if (L.nn) {
        check(copy_on_stream(L.host_nn_partials.data(), L.nn_partials.ptr, L.host_nn_partials.size() * sizeof(BlockStat),
                             hipMemcpyDeviceToHost),
              "network partials copy");
        check(copy_on_stream(L.host_nn_counts.data(), L.nn_counts.ptr, L.dates * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost),
              "network counts copy");
        check(copy_on_stream(&nn_in_sample, L.nn_in_sample.ptr, sizeof(double), hipMemcpyDeviceToHost),
              "network in-sample copy");
        pe_result network{};
        assemble(p, L.host_nn_partials, L.chunks, L.scrambles, 1, 0, 1, &network);
        network_est = network.estimates[0];
        fallback = judge.estimates[0].value > network.estimates[0].value;
        if (!fallback) *out = network;
    }
    const std::vector<unsigned long long>& counts = (L.nn && !fallback) ? L.host_nn_counts : L.host_counts;
    out->time_steps = L.n;
    out->path_steps = out->evaluated_paths * L.n;
    out->sobol_dimension = a.value.rqmc ? 2u * L.n : 0u;
    out->device_ms = train_ms + value_ms + nn_train_ms + nn_value_ms;
    g_kernel_ms = out->device_ms;
    g_kernel_launches = 3 + 6 * (L.dates > 1 ? L.dates - 1 : 0);
    if (!report) return;
    pe_lsm_report& r = *report;
    const pe_engine& te = p.req.lsm.training_engine;
    r.in_sample_value = L.nn ? nn_in_sample : in_sample;
    r.training_ms = train_ms + nn_train_ms;
    r.valuation_ms = value_ms + nn_value_ms;
    r.network_price = r.judge_price = pe_estimate{};
    r.judge_in_sample_value = 0.0;
    r.fallback_used = 0;
    r.network_training_ms = r.judge_training_ms = r.network_valuation_ms = r.judge_valuation_ms = 0.0;
    r.network_rows_trained = 0;
    if (L.nn) {
        r.network_price = network_est;
        r.judge_price = judge.estimates[0];
        r.judge_in_sample_value = in_sample;
        r.fallback_used = fallback ? 1 : 0;
        r.network_training_ms = path_ms + nn_train_ms;
        r.judge_training_ms = train_ms - path_ms;
        r.network_valuation_ms = nn_value_ms;
        r.judge_valuation_ms = value_ms;
        r.network_rows_trained = L.rows_trained;
    }
    r.path_ms = path_ms;
    r.exercise_dates = L.dates;
    r.training_random_domain = te.kind == PE_ENGINE_RQMC ? PE_DOMAIN_RQMC_SCRAMBLE : PE_DOMAIN_LSM_TRAIN;
    r.valuation_random_domain = a.value.rqmc ? PE_DOMAIN_RQMC_SCRAMBLE : PE_DOMAIN_VALUATION;
    r.training_sampling_units = te.kind == PE_ENGINE_RQMC ? te.scrambles : te.units;
    r.training_trajectories = a.train_paths;
    r.valuation_sampling_units = a.value.rqmc ? L.scrambles : a.value.units;
    r.valuation_trajectories = vtraj;
    r.training_has_checksums = p.train_diag.has_checksums;
    std::memcpy(r.training_direction_checksum, p.train_diag.direction_checksum, 32);
    std::memcpy(r.training_scramble_checksum, p.train_diag.scramble_checksum, 32);
    r.valuation_has_checksums = p.diag.has_checksums;
    std::memcpy(r.valuation_direction_checksum, p.diag.direction_checksum, 32);
    std::memcpy(r.valuation_scramble_checksum, p.diag.scramble_checksum, 32);
    r.policy_feature_count = 2;
    r.policy_max_degree = p.req.lsm.max_degree;
    r.itm_abs_tolerance = p.req.lsm.itm_abs_tolerance;
    r.abs_rank_tolerance = p.req.lsm.abs_rank_tolerance;
    r.rel_rank_tolerance = p.req.lsm.rel_rank_tolerance;
    r.max_matrix_elements = p.req.lsm.max_matrix_elements;
    if (want_stops) {
        const uint64_t k = std::min<uint64_t>(r.stopping_capacity, vtraj);
        check(copy_on_stream(r.stopping_indices, L.stops.ptr, k * sizeof(uint32_t), hipMemcpyDeviceToHost),
              "stopping indices copy");
    }
    for (uint32_t d = 0; d < L.dates && d < r.capacity; ++d) {
        if (r.exercise_counts) r.exercise_counts[d] = counts[d];
        if (r.exercise_probabilities)
            r.exercise_probabilities[d] = static_cast<double>(counts[d]) / static_cast<double>(vtraj);
        if (r.dividend_collisions) r.dividend_collisions[d] = 0;
        if (!r.decisions || d + 1 >= L.dates) continue;
        const RoughLsmDecision& s = L.host_decisions[d];
        pe_lsm_decision& o = r.decisions[d];
        std::memset(&o, 0, sizeof(o));
        o.feature_count = 2;
        o.continue_all = s.continue_all ? 1 : 0;
        o.itm_rows = s.itm_rows;
        if (s.continue_all) continue;
        o.rank = s.rank;
        o.active_columns = s.active;
        for (int k = 0; k < kRlCols; ++k) {
            o.coefficients[k] = s.coef[k];
            o.diagonal_abs[k] = s.diag[k];
            o.pivot_order[k] = s.pivot[k];
        }
        o.feature_mean = s.mean[0];
        o.feature_population_variance = s.variance[0];
        o.feature_scale = s.scale[0];
        o.feature_inactive = (s.inactive & 1u) ? 1 : 0;
        o.zero_scale_threshold = s.zero_threshold[0];
        o.variance_mean = s.mean[1];
        o.variance_population_variance = s.variance[1];
        o.variance_scale = s.scale[1];
        o.variance_inactive = (s.inactive & 2u) ? 1 : 0;
        o.variance_zero_scale_threshold = s.zero_threshold[1];
        o.rank_threshold = s.rank_threshold;
        o.residual_sum_squares = s.rss;
    }
}

void upload_lv_tables(HostPlan& p) {
    if (p.d_steps.ptr) return;
    p.d_steps.upload(p.steps);
    p.d_program.upload(p.program);
    p.d_grid.upload(p.lv_v);
    p.d_x.upload(p.lv_x);
    std::vector<float> g32(p.lv_v.begin(), p.lv_v.end()), x32(p.lv_x.begin(), p.lv_x.end());
    p.d_grid32.upload(g32);
    p.d_x32.upload(x32);
}

// Evaluates local-vol plans that share one simulation grid in one launch.
LvLaunch& prepare_lv(std::vector<HostPlan*>& plans, pe_backend backend) {
    HostPlan& p0 = *plans[0];
    std::vector<uint64_t> key;
    for (const HostPlan* q : plans) key.push_back(q->id);
    if (p0.lv_launch && p0.lv_launch->key == key && p0.lv_launch->backend == static_cast<int>(backend))
        return *p0.lv_launch;
    std::unique_ptr<LvLaunch> launch(new LvLaunch());
    LvLaunch& L = *launch;
    L.key = key;
    L.backend = static_cast<int>(backend);
    const pe_engine& e = p0.req.engine;
    if (plans.size() > static_cast<size_t>(kMaxProducts)) throw PlanError(PE_ERR_SIZE_LIMIT, "too many shared plans");
    for (HostPlan* q : plans) {
        if (!q->lv) throw PlanError(PE_ERR_INVALID_ARGUMENT, "shared evaluation needs local-vol plans");
        if (q->graph && plans.size() > 1)
            throw PlanError(PE_ERR_INVALID_ARGUMENT, "a graph product under local vol is evaluated on its own paths");
        if (q->risk && plans.size() > 1)
            throw PlanError(PE_ERR_INVALID_ARGUMENT, "shared evaluation needs price-only plans");
        if (q->nodes != p0.nodes || q->spot != p0.spot || q->node_forwards != p0.node_forwards ||
            q->lv_v != p0.lv_v || q->lv_x != p0.lv_x || q->lv_t != p0.lv_t ||
            std::memcmp(&q->req.engine, &e, sizeof(e)) != 0)
            throw PlanError(PE_ERR_INVALID_ARGUMENT, "plans do not share one simulated path set");
    }
    ensure_sobol(p0);
    upload_lv_tables(p0);
    const size_t nodes = p0.nodes.size();
    // Per-node fixing weights and the steps' node masks over all products.
    std::vector<double> node_weight(nodes * kMaxProducts, 0.0);
    std::vector<LvStep> steps = p0.steps;
    LvArgs& a = L.args;
    std::memset(&a, 0, sizeof(a));
    for (size_t k = 0; k < plans.size(); ++k) {
        const HostPlan& q = *plans[k];
        a.products[k] = q.product;
        for (size_t i = 0; i < q.obs_nodes.size(); ++i) {
            const uint32_t node = q.obs_nodes[i];
            node_weight[node * kMaxProducts + k] = q.obs_weight[i];
            // a graph reads every one of its observations, in order
            const bool needed = q.graph || q.obs_weight[i] != 0.0 || static_cast<int32_t>(node) == q.product.terminal_obs;
            if (!needed) continue;
            if (node > 0) steps[node - 1].node_mask |= 1 << k;
            else a.node0_mask |= 1u << k;
        }
    }
    // Fused instruction stream: bridge pushes and pops, or direct steps, in
    // an FP32 and an FP64 flavour.
    std::vector<LvOp32> ops32;
    std::vector<LvOp64> ops64;
    auto step_op = [&](uint32_t kind, int32_t step) {
        const LvStep& st = steps[static_cast<size_t>(step)];
        LvOp32 o32;
        LvOp64 o64;
        std::memset(&o32, 0, sizeof(o32));
        std::memset(&o64, 0, sizeof(o64));
        o32.code = o64.code = lv_code(kind, 0u, static_cast<uint32_t>(st.node_mask), static_cast<uint32_t>(st.t_cell));
        o32.index = o64.index = step;
        o32.inv_sqrt_dt = st.f_inv_sqrt_dt;
        o32.m_half_dt = st.m_half_dt;
        o32.sqrt_dt = st.f_sqrt_dt;
        o32.t_weight = st.f_t_weight;
        o32.next_forward = st.f_next_forward;
        o64.inv_sqrt_dt = st.inv_sqrt_dt;
        o64.dt = st.dt;
        o64.sqrt_dt = st.sqrt_dt;
        o64.ratio = st.ratio;
        o64.forward = st.forward;
        o64.t_weight = st.t_weight;
        ops32.push_back(o32);
        ops64.push_back(o64);
    };
    if (e.brownian_bridge) {
        for (const BridgeOp& b : p0.program) {
            if (b.push) {
                LvOp32 o32;
                LvOp64 o64;
                std::memset(&o32, 0, sizeof(o32));
                std::memset(&o64, 0, sizeof(o64));
                o32.code = o64.code = lv_code(kLvPush, b.terminal ? 1u : 0u, 0u, 0u);
                o32.index = o64.index = b.rank;
                o32.lw = b.f_left_weight;
                o32.rw = b.f_right_weight;
                o32.sd = b.f_sd;
                o64.lw = b.left_weight;
                o64.rw = b.right_weight;
                o64.sd = b.sd;
                ops32.push_back(o32);
                ops64.push_back(o64);
            } else {
                step_op(kLvPop, b.step);
            }
        }
    } else {
        for (size_t k = 0; k < steps.size(); ++k) step_op(kLvDirect, static_cast<int32_t>(k));
    }
    // the kernels read the pushes before each step up to that step: the
    // program must end with a step
    if (ops32.empty() || (ops32.back().code & 3u) == kLvPush)
        throw PlanError(PE_ERR_INVALID_ARGUMENT, "local-vol program must end with a step");
    for (size_t k = 0; k < plans.size(); ++k) {
        const ProductDesc& d = plans[k]->product;
        ProductHot& h = a.hot[k];
        h.terminal_node = d.terminal_obs;
        h.barrier = static_cast<float>(d.barrier);
        h.half_width = static_cast<float>(d.down ? -d.half_width : d.half_width);
        switch (d.kind) {
            case kAsian: h.mode = kModeAsian; break;
            case kBarrier: h.mode = d.smoothing ? kModeBarrierSmooth : (d.down ? kModeBarrierDown : kModeBarrierUp); break;
            case kLookback: h.mode = d.call ? kModeLookbackMax : kModeLookbackMin; break;
            default: h.mode = kModeTerminal;
        }
    }
    // Uniform-bin cell table for the mixed kernel.  The kernel maps the
    // clamped FP32 coordinate xx to bin b = trunc(fma(xx, inv_bin_w,
    // -x0 inv_bin_w)) in FP32; the same map is evaluated here on the FP32
    // nodes.  It is monotone
    // in xx, so bin_cell[b] = the number of interior nodes whose image lies
    // below b is a lower bound of xx's cell, and at most the next node (whose
    // image may equal b) remains to be compared.  The bin count is the
    // smallest power of two (up to 4096) at which no two interior nodes share
    // an image; the table covers every image up to that of the last node.
    std::vector<uint16_t> bin_cell;
    {
        const std::vector<double>& xs = p0.lv_x;
        const size_t nx = xs.size();
        std::vector<float> xf(nx);
        for (size_t i = 0; i < nx; ++i) xf[i] = static_cast<float>(xs[i]);
        const float x0f = xf.front(), xlf = xf.back();
        for (uint32_t m = 64; m <= 4096 && bin_cell.empty(); m *= 2) {
            const float inv = static_cast<float>(m / (xs.back() - xs.front()));
            const float off = -(x0f * inv);
            auto image = [&](float x) -> int32_t {
                const volatile float q = std::fmaf(x, inv, off);   // the kernel's single rounding
                return static_cast<int32_t>(q);
            };
            const int32_t last = image(xlf);
            bool ok = last >= 0 && last < 65536;
            std::vector<int32_t> img;
            for (size_t j = 1; ok && j + 1 < nx; ++j) {
                img.push_back(image(xf[j]));
                ok = img.size() < 2 || img[img.size() - 1] > img[img.size() - 2];
            }
            if (!ok) continue;
            std::vector<uint16_t> t(static_cast<size_t>(last) + 1);
            size_t below = 0;   // interior nodes with an image < b
            for (int32_t b = 0; b <= last; ++b) {
                while (below < img.size() && img[below] < b) ++below;
                t[static_cast<size_t>(b)] = static_cast<uint16_t>(below);
            }
            bin_cell.swap(t);
            a.inv_bin_w32 = inv;
            a.bin_off32 = off;
        }
        if (bin_cell.empty()) throw PlanError(PE_ERR_SIZE_LIMIT, "log-moneyness nodes too uneven for the cell table");
        a.bins = static_cast<uint32_t>(bin_cell.size());
        a.x0_32 = x0f;
        a.xl_32 = xlf;
    }
    DeviceBuffer& d_steps = L.d_steps;
    DeviceBuffer& d_node_weight = L.d_node_weight;
    DeviceBuffer& d_ops32 = L.d_ops32;
    DeviceBuffer& d_ops64 = L.d_ops64;
    L.d_bins.upload(bin_cell);
    a.bin_cell = L.d_bins.as<uint16_t>();
    d_steps.upload(steps);
    d_node_weight.upload(node_weight);
    a.n_ops = static_cast<uint32_t>(ops32.size());
    a.ops_stride = 0;
    if (e.kind == PE_ENGINE_RQMC && e.brownian_bridge) {
        // one copy per scramble, each push carrying its dimension's direction
        // words 0-5 of that scramble (the lane bits of the Sobol index)
        const std::vector<uint32_t> tab = build_sobol_table(e.seed, e.scrambles, p0.sobol_dims);
        const size_t n = ops32.size();
        std::vector<LvOp32> copies(n * e.scrambles);
        for (uint32_t sc = 0; sc < e.scrambles; ++sc)
            for (size_t k = 0; k < n; ++k) {
                LvOp32 o = ops32[k];
                if ((o.code & 3u) == kLvPush) {
                    const uint32_t dim = static_cast<uint32_t>(o.index);
                    if (dim >= p0.sobol_dims) throw PlanError(PE_ERR_INVALID_ARGUMENT, "bridge dimension out of range");
                    uint32_t words[12];
                    std::memcpy(words, &o, sizeof(o));
                    for (int b = 0; b < 6; ++b)
                        words[kLvPushDirWord + b] = tab[(static_cast<size_t>(sc) * p0.sobol_dims + dim) * 33u + b];
                    std::memcpy(&o, words, sizeof(o));
                }
                copies[static_cast<size_t>(sc) * n + k] = o;
            }
        ops32.swap(copies);
        a.ops_stride = static_cast<uint32_t>(n);
    }
    d_ops32.upload(ops32);
    d_ops64.upload(ops64);
    a.ops32 = d_ops32.as<LvOp32>();
    a.ops64 = d_ops64.as<LvOp64>();
    a.steps = d_steps.as<LvStep>();
    a.program = p0.d_program.as<BridgeOp>();
    a.grid = p0.d_grid.as<double>();
    a.grid32 = p0.d_grid32.as<float>();
    a.x_nodes = p0.d_x.as<double>();
    a.x_nodes32 = p0.d_x32.as<float>();
    a.sobol = p0.d_sobol.as<uint32_t>();
    a.node_weight = d_node_weight.as<double>();
    a.seed = e.seed;
    a.units = e.units;
    a.spot = p0.spot;
    a.n_steps = static_cast<uint32_t>(p0.steps.size());
    a.n_program = static_cast<uint32_t>(p0.program.size());
    a.nt = static_cast<uint32_t>(p0.lv_t.size());
    a.nx = static_cast<uint32_t>(p0.lv_x.size());
    a.n_products = static_cast<uint32_t>(plans.size());
    a.rqmc = e.kind == PE_ENGINE_RQMC;
    a.antithetic = e.antithetic ? 1u : 0u;
    a.bridge = e.brownian_bridge ? 1u : 0u;
    a.depth = p0.bridge_depth;
    a.sobol_dims = p0.sobol_dims;
    const dim3 grid = grid_for(e);
    const uint32_t comps = kMaxProducts;
    L.grid = grid;
    L.partial_count = static_cast<size_t>(grid.x) * grid.y * comps;
    L.partials.reserve(L.partial_count * sizeof(BlockStat));
    L.err.reserve(sizeof(int));
    check(hipHostMalloc(&L.host_partials, L.partial_count * sizeof(BlockStat), hipHostMallocDefault), "hipHostMalloc");
    check(hipHostMalloc(reinterpret_cast<void**>(&L.host_err), sizeof(int), hipHostMallocDefault), "hipHostMalloc");
    a.partials = L.partials.ptr;
    // LDS: grid, x nodes and inverse widths (Real), the per-wave Sobol words
    // (RQMC), the bin table (mixed), then the bridge stack (one slot per lane
    // per tree level).
    // (mixed: the cells' 16-byte records and the grid by column, columns of
    // nt | 1 words)
    const size_t real = backend == PE_BACKEND_GPU_FP64 ? sizeof(double) : sizeof(float);
    const size_t grid_bytes = backend == PE_BACKEND_GPU_FP64
                                  ? real * (a.nt * a.nx + 2u * a.nx)
                                  : 16u * a.nx + sizeof(float) * static_cast<size_t>(a.nx) * (a.nt | 1u);
    const size_t head = grid_bytes + (a.rqmc ? 4u * sizeof(uint32_t) * a.sobol_dims : 0u);
    const size_t bin_bytes = backend == PE_BACKEND_GPU_FP64 ? 0u : ((a.bins * sizeof(uint16_t) + 15u) & ~static_cast<size_t>(15u));
    L.lds = ((head + 15u) & ~static_cast<size_t>(15u)) + bin_bytes +
            static_cast<size_t>(std::max<uint32_t>(a.depth, 1u)) * kBlockThreads * real;
    if (p0.risk) {
        // lv_risk_kernel: bumped runtimes and the 2 sqrt(v) node tangents after the price layout
        L.risk = true;
        a.risk = (p0.request_vega ? kLvRiskVega : 0u) | (p0.request_gamma ? kLvRiskGamma : 0u) |
                 (p0.request_delta || p0.request_gamma ? kLvRiskBumps : 0u);
        a.spot_bump = p0.spot_bump;
        a.spot_dn = p0.spot - p0.spot_bump;
        a.spot_up = p0.spot + p0.spot_bump;
        a.scale_dn = a.spot_dn / p0.spot;
        a.scale_up = a.spot_up / p0.spot;
        std::vector<LvBumpStep> bumps = p0.lv_bumps;
        if (bumps.empty()) bumps.assign(1, LvBumpStep{1.0, 1.0, 1.0, 1.0});
        L.d_bumps.upload(bumps);
        a.bump_steps = L.d_bumps.as<LvBumpStep>();
        a.risk_lds_offset = static_cast<uint32_t>((L.lds + 15u) & ~static_cast<size_t>(15u));
        L.lds = a.risk_lds_offset + static_cast<size_t>(a.nt) * a.nx * real;
    }
    if (p0.graph) {
        // the graph's register program after the bridge stack: graph_regs
        // values per lane
        const GraphTape& gt = p0.graph_tape;
        if (gt.segs.size() != p0.obs_nodes.size() + 2)
            throw PlanError(PE_ERR_INVALID_ARGUMENT, "graph observations do not match the local-vol nodes");
        L.graph = true;
        L.d_graph_ops.upload(gt.ops.empty() ? std::vector<GraphOp>(1) : gt.ops);
        L.d_graph_segs.upload(gt.segs);
        L.d_graph_consts.upload(gt.consts.empty() ? std::vector<double>(1, 0.0) : gt.consts);
        a.graph_ops = L.d_graph_ops.as<GraphOp>();
        a.graph_segs = L.d_graph_segs.as<uint32_t>();
        a.graph_consts = L.d_graph_consts.as<double>();
        a.graph_output = gt.output;
        a.graph_regs = std::max<uint32_t>(1, gt.regs);
        a.graph_lds_offset = static_cast<uint32_t>((L.lds + 15u) & ~static_cast<size_t>(15u));
        L.lds = a.graph_lds_offset + static_cast<size_t>(a.graph_regs) * kBlockThreads * real;
    }
    if (L.lds > 160u * 1024u) throw PlanError(PE_ERR_SIZE_LIMIT, "local-vol grid and bridge stack exceed LDS");
    p0.lv_launch = std::move(launch);
    return *p0.lv_launch;
}

void evaluate_lv(std::vector<HostPlan*>& plans, pe_backend backend, pe_result* outs) {
    LvLaunch& L = prepare_lv(plans, backend);
    const int no_error = 0x7fffffff;
    *L.host_err = no_error;
    check(hipMemcpyAsync(L.err.ptr, L.host_err, sizeof(int), hipMemcpyHostToDevice, t_stream), "error word init");
    // launch, the copies back and one synchronisation; the kernel's time comes
    // from the events around the launch alone
    CallEvents& ev = call_events();
    check(hipEventRecord(ev.start, t_stream), "hipEventRecord");
    {
        if (L.risk)
            check(backend == PE_BACKEND_GPU_FP64 ? launch_lv_risk_f64(L.args, L.grid, L.lds, L.err.as<int>(), t_stream)
                                                 : launch_lv_risk_f32(L.args, L.grid, L.lds, L.err.as<int>(), t_stream),
                  "lv risk kernel launch");
        else if (L.graph)
            check(backend == PE_BACKEND_GPU_FP64 ? launch_lv_graph_f64(L.args, L.grid, L.lds, L.err.as<int>(), t_stream)
                                                 : launch_lv_graph_f32(L.args, L.grid, L.lds, L.err.as<int>(), t_stream),
                  "lv graph kernel launch");
        else
            check(backend == PE_BACKEND_GPU_FP64 ? launch_lv_f64(L.args, L.grid, L.lds, L.err.as<int>(), t_stream)
                                                 : launch_lv_f32(L.args, L.grid, L.lds, L.err.as<int>(), t_stream),
                  "lv kernel launch");
    }
    check(hipEventRecord(ev.stop, t_stream), "hipEventRecord");
    check(hipMemcpyAsync(L.host_partials, L.partials.ptr, L.partial_count * sizeof(BlockStat), hipMemcpyDeviceToHost,
                         t_stream), "partials copy");
    check(hipMemcpyAsync(L.host_err, L.err.ptr, sizeof(int), hipMemcpyDeviceToHost, t_stream), "error word copy");
    check(hipStreamSynchronize(t_stream), "stream synchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ev.start, ev.stop), "hipEventElapsedTime");
    const int failing = *L.host_err;
    if (failing != no_error) {
        for (size_t k = 0; k < plans.size(); ++k) outs[k].failing_unit = static_cast<uint64_t>(failing);
        throw PlanError(PE_ERR_NONFINITE_STATE,
                        "local-vol Log-Euler produced a non-finite state at sampling unit " + std::to_string(failing));
    }
    const BlockStat* partials = static_cast<const BlockStat*>(L.host_partials);
    for (size_t k = 0; k < plans.size(); ++k) {
        std::memset(&outs[k], 0, sizeof(pe_result));
        if (L.risk) {
            // slots: price, Delta, Vega, Gamma; the bump-validation components
            // repeat Delta and Gamma (pathwise_local_vol.rs:346-355)
            assemble(*plans[k], partials, L.grid.x, L.grid.y, kMaxProducts, 0, 4, &outs[k]);
            const HostPlan& q = *plans[k];
            if (q.request_delta || q.request_gamma) {
                outs[k].estimates[PE_C_BUMP_DELTA] = outs[k].estimates[PE_C_DELTA];
                outs[k].component_sum[PE_C_BUMP_DELTA] = outs[k].component_sum[PE_C_DELTA];
                outs[k].component_mean[PE_C_BUMP_DELTA] = outs[k].component_mean[PE_C_DELTA];
                outs[k].component_m2[PE_C_BUMP_DELTA] = outs[k].component_m2[PE_C_DELTA];
            }
            if (q.request_gamma) {
                outs[k].estimates[PE_C_BUMP_GAMMA] = outs[k].estimates[PE_C_GAMMA];
                outs[k].component_sum[PE_C_BUMP_GAMMA] = outs[k].component_sum[PE_C_GAMMA];
                outs[k].component_mean[PE_C_BUMP_GAMMA] = outs[k].component_mean[PE_C_GAMMA];
                outs[k].component_m2[PE_C_BUMP_GAMMA] = outs[k].component_m2[PE_C_GAMMA];
            }
        } else {
            assemble(*plans[k], partials, L.grid.x, L.grid.y, kMaxProducts, static_cast<uint32_t>(k), 1, &outs[k]);
        }
        outs[k].device_ms = ms;
    }
    g_kernel_ms = ms;
    g_kernel_launches = 1;
}

}  // namespace

// ------------------------------------------------------- early exercise --
// A set of American plans runs as three launches (training paths, one
// induction workgroup per trade, valuation items); all device state is cached
// on the set's first plan.
struct LsmLaunch {
    std::vector<uint64_t> key;
    int backend = 0;
    uint32_t basis = 0;
    std::vector<LsmTrade> trades;
    std::vector<LsmItem> train_items, value_items;
    std::vector<uint32_t> first_partial, chunks, scrambles;
    DeviceBuffer d_trades, d_dates, d_sobol, d_train_items, d_value_items, store, cash, decisions, counts, in_sample,
        partials, risk_partials, err, stops, gram;
    uint32_t gram_chunks = 0, max_dates = 0; // weighted-Laguerre pass partials per trade, most dates of a trade
    uint64_t n_trajectories = 0;             // valuation trajectories of all trades
    size_t n_decisions = 0, n_counts = 0, n_partials = 0;
    bool risk = false;                       // some plan asks for fixed-policy Greeks
    std::vector<BlockStat> host_partials, host_risk_partials;
    std::vector<LsmDecision> host_decisions;
    std::vector<unsigned long long> host_counts;
    std::vector<double> host_in_sample;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // start, training done, valuation done, paths done
    hipEvent_t mlp_ev[2] = {nullptr, nullptr};                 // network trained, network valued
    // network policy and its polynomial judge
    bool mlp = false;
    MlpLsmArgs mlp_args{};
    std::vector<LsmCuts> host_cuts;
    std::vector<BlockStat> host_judge_partials;
    std::vector<unsigned long long> host_judge_counts;
    std::vector<double> host_judge_in_sample;
    std::vector<unsigned long long> host_rows_trained;
    DeviceBuffer rows_trained, mlp_init, weights, adam_m, adam_v, mlp_rows, cuts, mlp_err, judge_cash, judge_partials, judge_counts,
        judge_in_sample;
    ~LsmLaunch() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : mlp_ev)
            if (e) (void)hipEventDestroy(e);
    }
};

void LsmLaunchDeleter::operator()(LsmLaunch* l) const { delete l; }

namespace {

// The network policy trains one network per trade (one workgroup each),
// from He-uniform hidden-layer weights drawn from SplitMix64 of the network
// seed and a zero output layer (every trade starts from the same vector;
// Adam state starts at zero for every evaluation).  The weighted-Laguerre judge shares the training paths
// and the valuation paths.
void prepare_mlp(LsmLaunch& L, std::vector<HostPlan*>& plans, size_t cash_bytes) {
    const HostPlan& p0 = *plans[0];
    const pe_lsm& l = p0.req.lsm;
    for (const HostPlan* q : plans) {
        const pe_lsm& o = q->req.lsm;
        if (o.mlp_hidden != l.mlp_hidden || o.mlp_epochs != l.mlp_epochs || o.mlp_batch != l.mlp_batch ||
            o.mlp_learning_rate != l.mlp_learning_rate || o.mlp_seed != l.mlp_seed)
            throw PlanError(PE_ERR_INVALID_ARGUMENT, "plans of one network-policy launch need one network setting");
        if (q->risk) throw PlanError(PE_ERR_UNSUPPORTED, "fixed-policy Greeks with the network policy");
        if (q->req.product.side != p0.req.product.side)
            throw PlanError(PE_ERR_INVALID_ARGUMENT, "plans of one network-policy launch need one option side");
    }
    if (l.mlp_hidden != 64) throw PlanError(PE_ERR_UNSUPPORTED, "MLP hidden width must be 64 on the device");
    if (l.mlp_epochs == 0 || !(l.mlp_learning_rate > 0.0)) throw PlanError(PE_ERR_INVALID_ARGUMENT, "MLP epochs / rate");
    const uint32_t B = l.mlp_batch;
    if (B == 0 || B % 128 != 0 || B > (1u << 20))
        throw PlanError(PE_ERR_INVALID_ARGUMENT, "MLP batch must be a positive multiple of 128, at most 2^20");
    L.mlp = true;
    for (hipEvent_t& e : L.mlp_ev) check(hipEventCreate(&e), "hipEventCreate");
    std::vector<float> init(kMlpParameterCount, 0.0f);
    uint64_t state = l.mlp_seed;
    auto uniform = [&]() {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return 2.0 * (static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    };
    for (int k = 0; k < 128; ++k) init[k] = static_cast<float>(uniform() * std::sqrt(6.0 / 2.0));
    for (int k = 192; k < 4288; ++k) init[k] = static_cast<float>(uniform() * std::sqrt(6.0 / 64.0));
    // the output layer starts at zero: the untrained network is the control
    // variate alone (continuation = the European value + the mean residual)
    L.mlp_init.upload(init);
    const size_t n = plans.size();
    const size_t stride = (kMlpParameterCount + 3) / 4 * 4;
    uint64_t rows = 0;
    for (const HostPlan* q : plans) rows += q->lsm_trade.train_paths;
    L.weights.reserve(n * stride * sizeof(float));
    L.adam_m.reserve(n * stride * sizeof(float));
    L.adam_v.reserve(n * stride * sizeof(float));
    L.mlp_rows.reserve(static_cast<size_t>(rows) * 4 * sizeof(float) + 16);
    L.cuts.reserve(std::max<size_t>(1, L.n_decisions) * sizeof(LsmCuts));
    L.mlp_err.reserve(sizeof(unsigned int));
    L.host_cuts.resize(L.n_decisions);
    L.judge_cash.reserve(cash_bytes + 256);
    L.judge_partials.reserve(std::max<size_t>(1, L.n_partials) * sizeof(BlockStat));
    L.judge_counts.reserve(L.n_counts * sizeof(unsigned long long));
    L.judge_in_sample.reserve(n * sizeof(double));
    L.host_judge_partials.resize(L.n_partials);
    L.host_judge_counts.resize(L.n_counts);
    L.host_judge_in_sample.resize(n);
    L.rows_trained.reserve(n * sizeof(unsigned long long));
    L.host_rows_trained.resize(n);
    MlpLsmArgs& m = L.mlp_args;
    m.init = L.mlp_init.as<float>();
    m.weights = L.weights.as<float>();
    m.adam_m = L.adam_m.as<float>();
    m.adam_v = L.adam_v.as<float>();
    m.rows = L.mlp_rows.as<float4>();
    m.cuts = L.cuts.as<LsmCuts>();
    m.error = L.mlp_err.as<unsigned int>();
    m.rows_trained = L.rows_trained.as<unsigned long long>();
    m.cash_bytes = cash_bytes;
    m.n_trades = static_cast<uint32_t>(n);
    m.batch = B;
    m.epochs = l.mlp_epochs;
    m.lr = static_cast<float>(l.mlp_learning_rate);
    m.beta1 = 0.9f;
    m.beta2 = 0.999f;
    m.eps = 1e-8f;
    const bool call = p0.req.product.side == PE_CALL;
    m.grid_lo = call ? 1.0f : 0.25f;
    m.grid_hi = call ? 4.0f : 1.0f;
}

LsmLaunch& prepare_lsm(std::vector<HostPlan*>& plans, pe_backend backend) {
    HostPlan& p0 = *plans[0];
    std::vector<uint64_t> key;
    for (const HostPlan* q : plans) key.push_back(q->id);
    if (p0.lsm_launch && p0.lsm_launch->key == key && p0.lsm_launch->backend == static_cast<int>(backend))
        return *p0.lsm_launch;
    auto L = std::make_unique<LsmLaunch>();
    L->key = key;
    L->backend = static_cast<int>(backend);
    // Training paths are FP64 in both modes for the polynomial policies: the
    // backward induction turns FP32 rounding of the stored spots into flipped
    // exercise decisions on near-boundary training paths (a few tenths of an
    // SE on some trades), so the mixed mode trains exactly as the FP64 mode
    // and runs only the valuation in FP32.  The network policy keeps its
    // FP32 store (its own tolerance is the polynomial policy's SE).
    const bool fp32_store = backend != PE_BACKEND_GPU_FP64 && plans[0]->lsm_trade.basis == kLsmMlp;
    const size_t real = fp32_store ? sizeof(float) : sizeof(double);
    std::vector<LsmDate> dates;
    std::vector<uint32_t> sobol;
    uint64_t store_elems = 0, cash_elems = 0;
    uint32_t partials = 0;
    L->basis = plans[0]->lsm_trade.basis;
    for (size_t ti = 0; ti < plans.size(); ++ti) {
        HostPlan& p = *plans[ti];
        if (!p.american) throw PlanError(PE_ERR_INVALID_ARGUMENT, "not an early-exercise plan");
        if (p.lsm_trade.basis != L->basis) throw PlanError(PE_ERR_INVALID_ARGUMENT, "one LSM method per launch");
        LsmTrade t = p.lsm_trade;
        const uint32_t n = t.n_dates;
        t.date_offset = static_cast<uint32_t>(dates.size());
        dates.insert(dates.end(), p.lsm_dates.begin(), p.lsm_dates.end());
        t.store_offset = store_elems;
        t.cash_offset = cash_elems;
        store_elems += t.train_paths * n;
        cash_elems += t.train_paths;
        t.decision_offset = static_cast<uint32_t>(L->n_decisions);
        t.count_offset = static_cast<uint32_t>(L->n_counts);
        L->n_decisions += n - 1;
        L->n_counts += n;
        t.stop_offset = L->n_trajectories;
        L->n_trajectories += t.value.units * t.value.scrambles * (t.value.antithetic ? 2u : 1u);
        for (LsmEngine* e : {&t.train, &t.value}) {
            if (!e->rqmc) continue;
            e->sobol_offset = static_cast<uint32_t>(sobol.size());
            const std::vector<uint32_t> tab = build_sobol_table(e->seed, e->scrambles, n);
            if (sobol.size() + tab.size() > 0xffffffffull) throw PlanError(PE_ERR_SIZE_LIMIT, "Sobol tables");
            sobol.insert(sobol.end(), tab.begin(), tab.end());
        }
        L->trades.push_back(t);
        L->gram_chunks = std::max<uint32_t>(
            L->gram_chunks, static_cast<uint32_t>((t.train_paths + kLsmPassPaths - 1) / kLsmPassPaths));
        L->max_dates = std::max(L->max_dates, n);
        auto add_items = [&](const LsmEngine& e, std::vector<LsmItem>& items, bool value) {
            const uint32_t ch = static_cast<uint32_t>((e.units + kItemUnits - 1) / kItemUnits);
            if (value) {
                L->first_partial.push_back(partials);
                L->chunks.push_back(ch);
                L->scrambles.push_back(e.scrambles);
            }
            for (uint32_t sc = 0; sc < e.scrambles; ++sc)
                for (uint32_t c = 0; c < ch; ++c) {
                    LsmItem it{};
                    it.trade = static_cast<uint32_t>(ti);
                    it.scramble = sc;
                    it.unit_begin = static_cast<uint64_t>(c) * kItemUnits;
                    it.unit_end = std::min<uint64_t>(it.unit_begin + kItemUnits, e.units);
                    it.partial = value ? partials++ : 0u;
                    items.push_back(it);
                }
        };
        add_items(t.train, L->train_items, false);
        add_items(t.value, L->value_items, true);
    }
    // longest work first
    auto by_cost = [&](std::vector<LsmItem>& items) {
        std::stable_sort(items.begin(), items.end(), [&](const LsmItem& x, const LsmItem& y) {
            return (x.unit_end - x.unit_begin) * L->trades[x.trade].n_dates >
                   (y.unit_end - y.unit_begin) * L->trades[y.trade].n_dates;
        });
        if (items.size() > 0x7fffffffu) throw PlanError(PE_ERR_SIZE_LIMIT, "too many work items");
    };
    by_cost(L->train_items);
    by_cost(L->value_items);
    L->n_partials = partials;
    L->d_trades.upload(L->trades);
    L->d_dates.upload(dates);
    if (!sobol.empty()) L->d_sobol.upload(sobol);
    L->d_train_items.upload(L->train_items);
    L->d_value_items.upload(L->value_items);
    L->store.reserve(static_cast<size_t>(store_elems) * real + 256);
    L->cash.reserve(static_cast<size_t>(cash_elems) * real + 256);
    L->decisions.reserve(std::max<size_t>(1, L->n_decisions) * sizeof(LsmDecision));
    L->counts.reserve(L->n_counts * sizeof(unsigned long long));
    L->in_sample.reserve(plans.size() * sizeof(double));
    L->partials.reserve(std::max<size_t>(1, L->n_partials) * sizeof(BlockStat));
    for (const HostPlan* q : plans) L->risk = L->risk || q->risk;
    if (L->risk) {
        L->risk_partials.reserve(std::max<size_t>(1, L->n_partials) * kLsmRiskComponents * sizeof(BlockStat));
        L->host_risk_partials.resize(L->n_partials * kLsmRiskComponents);
    }
    L->err.reserve(sizeof(unsigned long long));
    L->gram_chunks = std::max<uint32_t>(1, L->gram_chunks);
    if (static_cast<uint64_t>(plans.size()) * L->gram_chunks > 0x7fffffffull)
        throw PlanError(PE_ERR_SIZE_LIMIT, "too many training-path chunks");
    L->gram.reserve(plans.size() * L->gram_chunks * 16 * sizeof(double));
    L->host_partials.resize(L->n_partials);
    L->host_decisions.resize(L->n_decisions);
    L->host_counts.resize(L->n_counts);
    L->host_in_sample.resize(plans.size());
    for (hipEvent_t& e : L->ev) check(hipEventCreate(&e), "hipEventCreate");
    if (L->basis == kLsmMlp) prepare_mlp(*L, plans, static_cast<size_t>(cash_elems) * real);
    p0.lsm_launch.reset(L.release());
    return *p0.lsm_launch;
}

// Andersen-Broadie bounds (AndersenBroadiePlan::evaluate,
// crates/pricing/src/engine/risk/dual.rs:219-296): the LSM training launch,
// then the inner-rollout and outer kernels.
void evaluate_lsm_dual(HostPlan* plan, uint32_t n2, uint32_t n3, uint64_t seed, pe_dual_result* out) {
    HostPlan& p = *plan;
    const pe_engine& ve = p.req.engine;
    const pe_engine& te = p.req.lsm.training_engine;
    if (!p.american) throw PlanError(PE_ERR_INVALID_ARGUMENT, "Andersen-Broadie needs an American plan");
    if (p.lv || p.rough) throw PlanError(PE_ERR_UNSUPPORTED, "Andersen-Broadie beyond Black-Scholes/Black-76");
    if (p.risk) throw PlanError(PE_ERR_UNSUPPORTED, "Greeks in Andersen-Broadie (use a fixed-policy risk request)");
    if (ve.kind != PE_ENGINE_PSEUDO || te.kind != PE_ENGINE_PSEUDO)
        throw PlanError(PE_ERR_UNSUPPORTED, "RQMC or mixed training/valuation engines in Andersen-Broadie");
    if (p.req.lsm.method == PE_LSM_MLP) throw PlanError(PE_ERR_UNSUPPORTED, "the network policy in Andersen-Broadie");
    if (n2 == 0 || n3 == 0) throw PlanError(PE_ERR_INVALID_ARGUMENT, "inner-path counts must be positive");
    const uint64_t lanes = ve.antithetic ? 2 : 1;
    const uint64_t trajectories = ve.units * lanes;
    const uint64_t stride = std::max(n2, n3);
    if (trajectories > UINT64_MAX / stride) throw PlanError(PE_ERR_SIZE_LIMIT, "Andersen-Broadie random-coordinate count");
    const uint32_t n = p.lsm_trade.n_dates;
    if (static_cast<uint64_t>(n) * n > 0xffffffffull) throw PlanError(PE_ERR_SIZE_LIMIT, "Andersen-Broadie dimensions");
    const uint64_t pairs = trajectories * (n - 1);
    if (pairs > 0x7fffffffull) throw PlanError(PE_ERR_SIZE_LIMIT, "too many outer trajectory x date rollouts");
    std::vector<HostPlan*> plans{plan};
    LsmLaunch& L = prepare_lsm(plans, PE_BACKEND_GPU_FP64);
    const unsigned long long none = ~0ull;
    check(hipMemcpyAsync(L.err.ptr, &none, sizeof(none), hipMemcpyHostToDevice, t_stream), "error word");
    LsmArgs a{};
    a.trades = L.d_trades.as<LsmTrade>();
    a.dates = L.d_dates.as<LsmDate>();
    a.sobol = L.d_sobol.as<uint32_t>();
    a.store = L.store.ptr;
    a.cash = L.cash.ptr;
    a.decisions = L.decisions.as<LsmDecision>();
    a.counts = L.counts.as<unsigned long long>();
    a.in_sample = L.in_sample.as<double>();
    a.partials = L.partials.ptr;
    a.error = L.err.as<unsigned long long>();
    a.n_trades = 1;
    a.gram = L.gram.as<double>();
    a.gram_chunks = L.gram_chunks;
    a.max_dates = L.max_dates;
    LsmArgs train = a;
    train.items = L.d_train_items.as<LsmItem>();
    train.n_items = static_cast<uint32_t>(L.train_items.size());
    DeviceBuffer cont, partials;
    cont.reserve(std::max<uint64_t>(1, pairs) * sizeof(double));
    const uint32_t blocks = static_cast<uint32_t>((ve.units + kBlockThreads - 1) / kBlockThreads);
    partials.reserve(static_cast<size_t>(blocks) * 3 * sizeof(BlockStat));
    LsmDualArgs da{};
    da.lsm = a;
    da.cont = cont.as<double>();
    da.partials = partials.ptr;
    da.inner_seed = seed;
    da.stride = stride;
    da.n2 = n2;
    da.n3 = n3;
    da.lanes = static_cast<uint32_t>(lanes);
    const float ms = timed([&] {
        check(launch_lsm_f64(train, train.n_items, train, 0, L.basis, false, t_stream, nullptr, nullptr),
              "LSM training launch");
        check(launch_lsm_dual_f64(da, pairs, blocks, L.basis, t_stream), "Andersen-Broadie launch");
    });
    unsigned long long err = none;
    check(copy_on_stream(&err, L.err.ptr, sizeof(err), hipMemcpyDeviceToHost), "error copy");
    if (err != none) throw PlanError(PE_ERR_NONFINITE_STATE, "non-finite early-exercise training path");
    std::vector<BlockStat> host(static_cast<size_t>(blocks) * 3);
    check(copy_on_stream(host.data(), partials.ptr, host.size() * sizeof(BlockStat), hipMemcpyDeviceToHost),
          "partials copy");
    pe_result r{};
    assemble(p, host, blocks, 1, 3, 0, 3, &r);
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(r.estimates[c].value)) throw PlanError(PE_ERR_NONFINITE_STATE, "non-finite dual bound");
    std::memset(out, 0, sizeof(*out));
    out->lower_bound = r.estimates[0];
    out->upper_bound = r.estimates[1];
    out->duality_gap = r.estimates[2];
    out->price_confidence_interval_95[0] = r.estimates[0].confidence_lower;
    out->price_confidence_interval_95[1] = r.estimates[1].confidence_upper;
    out->outer_trajectories = trajectories;
    out->exercise_date_count = n;
    out->device_ms = ms;
    g_kernel_ms = ms;
}

void evaluate_lsm(std::vector<HostPlan*>& plans, pe_backend backend, pe_result* outs, pe_lsm_report* reports) {
    LsmLaunch& L = prepare_lsm(plans, backend);
    const unsigned long long none = ~0ull;
    check(hipMemcpyAsync(L.err.ptr, &none, sizeof(none), hipMemcpyHostToDevice, t_stream), "error word");
    check(hipMemsetAsync(L.counts.ptr, 0, L.n_counts * sizeof(unsigned long long), t_stream), "counts");
    LsmArgs a{};
    a.trades = L.d_trades.as<LsmTrade>();
    a.dates = L.d_dates.as<LsmDate>();
    a.sobol = L.d_sobol.as<uint32_t>();
    a.store = L.store.ptr;
    a.cash = L.cash.ptr;
    a.decisions = L.decisions.as<LsmDecision>();
    a.counts = L.counts.as<unsigned long long>();
    a.in_sample = L.in_sample.as<double>();
    a.partials = L.partials.ptr;
    a.risk_partials = L.risk ? L.risk_partials.ptr : nullptr;
    a.error = L.err.as<unsigned long long>();
    a.n_trades = static_cast<uint32_t>(plans.size());
    a.gram = L.gram.as<double>();
    a.gram_chunks = L.gram_chunks;
    a.max_dates = L.max_dates;
    bool want_stops = false;
    for (size_t ti = 0; reports && ti < plans.size(); ++ti)
        want_stops = want_stops || (reports[ti].stopping_indices && reports[ti].stopping_capacity);
    if (want_stops) {
        L.stops.reserve(std::max<uint64_t>(1, L.n_trajectories) * sizeof(uint32_t));
        a.stops = L.stops.as<uint32_t>();
    }
    LsmArgs train = a, value = a;
    train.stops = nullptr;
    train.items = L.d_train_items.as<LsmItem>();
    train.n_items = static_cast<uint32_t>(L.train_items.size());
    value.items = L.d_value_items.as<LsmItem>();
    value.n_items = static_cast<uint32_t>(L.value_items.size());
    LsmArgs judge = train, judge_value = value;
    if (L.mlp) {
        train.cuts = value.cuts = L.cuts.as<LsmCuts>();
        L.mlp_args.lsm = train;
        judge.cash = L.judge_cash.ptr;
        judge.in_sample = L.judge_in_sample.as<double>();
        judge_value.partials = L.judge_partials.ptr;
        judge_value.counts = L.judge_counts.as<unsigned long long>();
        judge_value.stops = nullptr;
        judge_value.risk_partials = nullptr;
        check(hipMemsetAsync(L.mlp_err.ptr, 0, sizeof(unsigned int), t_stream), "MLP error word");
        check(hipMemsetAsync(L.cuts.ptr, 0, std::max<size_t>(1, L.n_decisions) * sizeof(LsmCuts), t_stream), "cuts");
        check(hipMemsetAsync(L.judge_counts.ptr, 0, L.n_counts * sizeof(unsigned long long), t_stream), "counts");
    }
    check(hipEventRecord(L.ev[0], t_stream), "hipEventRecord");
    // network launch marks: paths, network trained, judge trained, network valued
    const hipEvent_t mlp_marks[4] = {L.ev[3], L.mlp_ev[0], L.ev[1], L.mlp_ev[1]};
    if (L.mlp)
        check(backend == PE_BACKEND_GPU_FP64
                  ? launch_lsm_mlp_f64(train, train.n_items, L.mlp_args, judge, value, judge_value, value.n_items,
                                       L.risk, t_stream, mlp_marks)
                  : launch_lsm_mlp_f32(train, train.n_items, L.mlp_args, judge, value, judge_value, value.n_items,
                                       L.risk, t_stream, mlp_marks),
              "MLP LSM launch");
    else if (backend == PE_BACKEND_GPU_FP64)
        check(launch_lsm_f64(train, train.n_items, value, value.n_items, L.basis, L.risk, t_stream, L.ev[3], L.ev[1]),
              "LSM kernel launch");
    else {
        // FP64 training (bit-identical decisions to the FP64 mode), FP32 valuation
        check(launch_lsm_f64(train, train.n_items, value, 0, L.basis, false, t_stream, L.ev[3], L.ev[1]),
              "LSM training launch");
        check(launch_lsm_value_f32(value, value.n_items, L.basis, L.risk, t_stream), "LSM valuation launch");
    }
    check(hipEventRecord(L.ev[2], t_stream), "hipEventRecord");
    check(hipEventSynchronize(L.ev[2]), "hipEventSynchronize");
    float train_ms = 0.0f, value_ms = 0.0f, path_ms = 0.0f;
    check(hipEventElapsedTime(&train_ms, L.ev[0], L.ev[1]), "hipEventElapsedTime");
    check(hipEventElapsedTime(&path_ms, L.ev[0], L.ev[3]), "hipEventElapsedTime");
    check(hipEventElapsedTime(&value_ms, L.ev[1], L.ev[2]), "hipEventElapsedTime");
    // network policy: its own training (paths + network) and valuation, and the judge's
    float nn_train_ms = 0.0f, judge_train_ms = 0.0f, nn_value_ms = 0.0f, judge_value_ms = 0.0f;
    if (L.mlp) {
        check(hipEventElapsedTime(&nn_train_ms, L.ev[0], L.mlp_ev[0]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&judge_train_ms, L.mlp_ev[0], L.ev[1]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&nn_value_ms, L.ev[1], L.mlp_ev[1]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&judge_value_ms, L.mlp_ev[1], L.ev[2]), "hipEventElapsedTime");
    }
    check(copy_on_stream(L.host_partials.data(), L.partials.ptr, L.n_partials * sizeof(BlockStat), hipMemcpyDeviceToHost),
          "partials copy");
    if (L.risk)
        check(copy_on_stream(L.host_risk_partials.data(), L.risk_partials.ptr,
                        L.host_risk_partials.size() * sizeof(BlockStat), hipMemcpyDeviceToHost),
              "risk partials copy");
    if (L.n_decisions)
        check(copy_on_stream(L.host_decisions.data(), L.decisions.ptr, L.n_decisions * sizeof(LsmDecision),
                        hipMemcpyDeviceToHost),
              "decisions copy");
    check(copy_on_stream(L.host_counts.data(), L.counts.ptr, L.n_counts * sizeof(unsigned long long),
                    hipMemcpyDeviceToHost),
          "counts copy");
    check(copy_on_stream(L.host_in_sample.data(), L.in_sample.ptr, plans.size() * sizeof(double), hipMemcpyDeviceToHost),
          "in-sample copy");
    unsigned long long err = none;
    check(copy_on_stream(&err, L.err.ptr, sizeof(err), hipMemcpyDeviceToHost), "error copy");
    if (err != none) {
        const uint64_t trade = err >> 40, unit = err & ((1ull << 40) - 1);
        throw PlanError(PE_ERR_NONFINITE_STATE, "trade " + std::to_string(trade) + ", sampling unit " +
                                                    std::to_string(unit) + ": non-finite early-exercise path");
    }
    if (L.mlp) {
        unsigned int merr = 0;
        check(copy_on_stream(&merr, L.mlp_err.ptr, sizeof(merr), hipMemcpyDeviceToHost), "MLP error copy");
        if (merr) throw PlanError(PE_ERR_SIZE_LIMIT, "an MLP exercise region needs more moneyness cuts than the grid allows");
        if (L.n_decisions)
            check(copy_on_stream(L.host_cuts.data(), L.cuts.ptr, L.n_decisions * sizeof(LsmCuts), hipMemcpyDeviceToHost),
                  "cuts copy");
        check(copy_on_stream(L.host_judge_partials.data(), L.judge_partials.ptr, L.n_partials * sizeof(BlockStat),
                             hipMemcpyDeviceToHost),
              "judge partials copy");
        check(copy_on_stream(L.host_judge_counts.data(), L.judge_counts.ptr, L.n_counts * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost),
              "judge counts copy");
        check(copy_on_stream(L.host_judge_in_sample.data(), L.judge_in_sample.ptr, plans.size() * sizeof(double),
                             hipMemcpyDeviceToHost),
              "judge in-sample copy");
        check(copy_on_stream(L.host_rows_trained.data(), L.rows_trained.ptr, plans.size() * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost),
              "rows-trained copy");
    }
    for (size_t ti = 0; ti < plans.size(); ++ti) {
        const HostPlan& p = *plans[ti];
        const LsmTrade& t = L.trades[ti];
        const uint32_t per = L.scrambles[ti] * L.chunks[ti];
        std::vector<BlockStat> mine(L.host_partials.begin() + L.first_partial[ti],
                                    L.host_partials.begin() + L.first_partial[ti] + per);
        pe_result* out = outs + ti;
        std::memset(out, 0, sizeof(*out));
        assemble(p, mine, L.chunks[ti], L.scrambles[ti], 1, 0, 1, out);
        // network policy: the polynomial judge's out-of-sample value on the
        // same valuation paths; the higher of the two is the price
        pe_result judge{};
        bool fallback = false;
        pe_estimate network = out->estimates[0];
        if (L.mlp) {
            std::vector<BlockStat> jm(L.host_judge_partials.begin() + L.first_partial[ti],
                                      L.host_judge_partials.begin() + L.first_partial[ti] + per);
            assemble(p, jm, L.chunks[ti], L.scrambles[ti], 1, 0, 1, &judge);
            fallback = judge.estimates[0].value > out->estimates[0].value;
            if (fallback) *out = judge;
        }
        const std::vector<unsigned long long>& counts = fallback ? L.host_judge_counts : L.host_counts;
        if (p.risk) {
            // fixed-policy Greeks: components 1..9 from the valuation kernel's risk partials
            // (execute_early_exercise, crates/pricing/src/engine/risk/early_exercise.rs:115-170)
            const size_t first = static_cast<size_t>(L.first_partial[ti]) * kLsmRiskComponents;
            std::vector<BlockStat> rmine(L.host_risk_partials.begin() + first,
                                         L.host_risk_partials.begin() + first + static_cast<size_t>(per) * kLsmRiskComponents);
            pe_result r{};
            assemble(p, rmine, L.chunks[ti], L.scrambles[ti], kLsmRiskComponents, 0, kLsmRiskComponents, &r);
            for (int c = 0; c < kLsmRiskComponents; ++c) {
                out->estimates[1 + c] = r.estimates[c];
                out->component_sum[1 + c] = r.component_sum[c];
                out->component_mean[1 + c] = r.component_mean[c];
                out->component_m2[1 + c] = r.component_m2[c];
            }
        }
        out->device_ms = train_ms + value_ms;
        if (!reports) continue;
        pe_lsm_report& r = reports[ti];
        r.in_sample_value = L.host_in_sample[ti];
        r.training_ms = train_ms;
        r.valuation_ms = value_ms;
        r.path_ms = path_ms;
        r.exercise_dates = t.n_dates;
        // EarlyExerciseDiagnostics (risk/early_exercise.rs:109-113, 190-210)
        const uint64_t vtraj = t.value.units * t.value.scrambles * (t.value.antithetic ? 2u : 1u);
        // RQMC engines draw from the scramble domain (the reported RandomDomain)
        r.training_random_domain = t.train.rqmc ? PE_DOMAIN_RQMC_SCRAMBLE : PE_DOMAIN_LSM_TRAIN;
        r.valuation_random_domain = t.value.rqmc ? PE_DOMAIN_RQMC_SCRAMBLE : PE_DOMAIN_VALUATION;
        // sampling units are the independent units: pseudo-MC units, or RQMC scrambles
        r.training_sampling_units = t.train.rqmc ? t.train.scrambles : t.train.units;
        r.training_trajectories = t.train_paths;
        r.valuation_sampling_units = t.value.rqmc ? t.value.scrambles : t.value.units;
        r.valuation_trajectories = vtraj;
        r.training_has_checksums = p.train_diag.has_checksums;
        std::memcpy(r.training_direction_checksum, p.train_diag.direction_checksum, 32);
        std::memcpy(r.training_scramble_checksum, p.train_diag.scramble_checksum, 32);
        r.valuation_has_checksums = p.diag.has_checksums;
        std::memcpy(r.valuation_direction_checksum, p.diag.direction_checksum, 32);
        std::memcpy(r.valuation_scramble_checksum, p.diag.scramble_checksum, 32);
        r.policy_feature_count = 1;
        r.policy_max_degree = p.req.lsm.max_degree;
        r.itm_abs_tolerance = p.req.lsm.itm_abs_tolerance;
        r.abs_rank_tolerance = p.req.lsm.abs_rank_tolerance;
        r.rel_rank_tolerance = p.req.lsm.rel_rank_tolerance;
        r.max_matrix_elements = p.req.lsm.max_matrix_elements;
        r.network_price = r.judge_price = pe_estimate{};
        r.judge_in_sample_value = 0.0;
        r.fallback_used = 0;
        r.network_training_ms = r.judge_training_ms = r.network_valuation_ms = r.judge_valuation_ms = 0.0;
        r.network_rows_trained = 0;
        if (L.mlp) {
            r.network_price = network;
            r.judge_price = judge.estimates[0];
            r.judge_in_sample_value = L.host_judge_in_sample[ti];
            r.fallback_used = fallback ? 1 : 0;
            r.network_training_ms = nn_train_ms;
            r.judge_training_ms = judge_train_ms;
            r.network_valuation_ms = nn_value_ms;
            r.judge_valuation_ms = judge_value_ms;
            r.network_rows_trained = L.host_rows_trained[ti];
        }
        if (r.stopping_indices && r.stopping_capacity) {
            const uint64_t k = std::min<uint64_t>(r.stopping_capacity, vtraj);
            check(copy_on_stream(r.stopping_indices, L.stops.as<uint32_t>() + t.stop_offset, k * sizeof(uint32_t),
                            hipMemcpyDeviceToHost),
                  "stopping indices copy");
        }
        for (uint32_t d = 0; d < t.n_dates && d < r.capacity; ++d) {
            if (r.exercise_probabilities)
                r.exercise_probabilities[d] =
                    static_cast<double>(counts[t.count_offset + d]) / static_cast<double>(vtraj);
            if (r.dividend_collisions) r.dividend_collisions[d] = p.lsm_collisions[d];
            if (r.exercise_counts) r.exercise_counts[d] = counts[t.count_offset + d];
            if (!r.decisions || d + 1 >= t.n_dates) continue;
            const LsmDecision& s = L.host_decisions[t.decision_offset + d];
            pe_lsm_decision& o = r.decisions[d];
            std::memset(&o, 0, sizeof(o));
            if (L.mlp) {
                const LsmCuts& c = L.host_cuts[t.decision_offset + d];
                o.region_left = static_cast<int32_t>(c.left);
                o.region_cut_count = c.count;
                for (uint32_t k = 0; k < c.count && k < 64; ++k) o.region_cuts[k] = c.cut[k];
            }
            o.continue_all = s.continue_all ? 1 : 0;
            o.itm_rows = s.itm_rows;
            if (s.continue_all) continue;
            o.rank = s.rank;
            o.feature_count = 1;
            o.active_columns = (L.basis == kLsmMonomial && s.inactive) ? 1u : t.columns;
            for (int k = 0; k < 4; ++k) {
                o.coefficients[k] = s.coef[k];
                o.diagonal_abs[k] = s.diag[k];
                o.pivot_order[k] = s.pivot[k];
            }
            o.feature_mean = s.mean;
            o.feature_population_variance = s.variance;
            o.feature_scale = s.scale;
            o.feature_inactive = s.inactive ? 1 : 0;
            o.rank_threshold = s.rank_threshold;
            o.residual_sum_squares = s.rss;
            o.zero_scale_threshold = s.zero_threshold;
        }
    }
    g_kernel_ms = train_ms + value_ms;
    g_kernel_launches = 3;
}

template <typename F>
pe_status guarded(F&& f) {
    try {
        f();
        return PE_OK;
    } catch (const PlanError& e) {
        return fail(e.status, e.what());
    } catch (const std::bad_alloc&) {
        return fail(PE_ERR_ALLOCATION, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(PE_ERR_INVALID_ARGUMENT, e.what());
    }
}

}  // namespace

hipStream_t current_stream() { return t_stream; }

}  // namespace pe

using namespace pe;

extern "C" {

int32_t pe_abi_version(void) { return PE_ABI_VERSION; }

const char* pe_last_error(void) { return g_error.c_str(); }

pe_status pe_init(int32_t device) {
    return guarded([&] {
        g_device = device;
        ensure_init();
    });
}

pe_status pe_plan_compile(const pe_request* request, const pe_exec_policy* policy, pe_plan** out) {
    if (!request || !out) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    return guarded([&] {
        std::unique_ptr<pe_plan> plan(new pe_plan());
        compile_host_plan(plan->host, *request, policy);
        *out = plan.release();
    });
}

void pe_plan_free(pe_plan* plan) { delete plan; }

pe_status pe_plan_diagnostics(const pe_plan* plan, pe_mc_diagnostics* out) {
    if (!plan || !out) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    *out = plan->host.diag;
    return PE_OK;
}

pe_status pe_blake3_256(const void* data, size_t bytes, uint8_t out[32]) {
    if ((!data && bytes) || !out) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    Blake3 b;
    b.update(data, bytes);
    b.finalize(out);
    return PE_OK;
}

pe_status pe_plan_evaluate(const pe_plan* plan, pe_backend backend, pe_result* out) {
    if (!plan || !out) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    if (backend != PE_BACKEND_GPU_FP64 && backend != PE_BACKEND_GPU_MIXED) return fail(PE_ERR_INVALID_ARGUMENT, "backend");
    return guarded([&] {
        ensure_init();
        HostPlan& p = const_cast<pe_plan*>(plan)->host;
        CallStream cs;
        PlanLocks locks({&p});
        if (p.rough && p.american) {
            evaluate_rough_lsm(p, backend, out, nullptr);
        } else if (p.rough) {
            std::vector<HostPlan*> plans{&p};
            evaluate_rough(plans, backend, out);
        } else if (p.american) {
            std::vector<HostPlan*> plans{&p};
            evaluate_lsm(plans, backend, out, nullptr);
        } else if (p.lv) {
            std::vector<HostPlan*> plans{&p};
            evaluate_lv(plans, backend, out);
        } else {
            std::vector<HostPlan*> plans{&p};
            evaluate_bs_batch(plans, backend, out);
        }
    });
}

pe_status pe_plan_evaluate_noalloc(const pe_plan* plan, pe_backend backend, pe_result* out) {
    if (!plan || !out) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    if (backend != PE_BACKEND_GPU_FP64 && backend != PE_BACKEND_GPU_MIXED) return fail(PE_ERR_INVALID_ARGUMENT, "backend");
    return guarded([&] {
        ensure_init();
        CallStream cs;
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        PlanLocks locks({&const_cast<pe_plan*>(plan)->host});
        evaluate_bs_noalloc(const_cast<pe_plan*>(plan)->host, backend, out);
    });
}

pe_status pe_plans_evaluate_shared(const pe_plan* const* plans, uint32_t count, pe_backend backend, pe_result* out) {
    if (!plans || !out || count == 0) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    if (backend != PE_BACKEND_GPU_FP64 && backend != PE_BACKEND_GPU_MIXED) return fail(PE_ERR_INVALID_ARGUMENT, "backend");
    return guarded([&] {
        ensure_init();
        std::vector<HostPlan*> hp;
        for (uint32_t i = 0; i < count; ++i) {
            if (!plans[i]) throw PlanError(PE_ERR_INVALID_ARGUMENT, "null plan");
            hp.push_back(&const_cast<pe_plan*>(plans[i])->host);
        }
        CallStream cs;
        PlanLocks locks(hp);
        for (const HostPlan* q : hp)
            if (q->rough && q->american)
                throw PlanError(PE_ERR_INVALID_ARGUMENT, "a lifted-Heston Bermudan is evaluated on its own paths");
        if (hp[0]->rough) evaluate_rough(hp, backend, out);
        else evaluate_lv(hp, backend, out);
    });
}

pe_status pe_lifted_heston_from_rough(double hurst, uint32_t factors, double ratio, double* weights_out,
                                      double* rates_out) {
    if (!weights_out || !rates_out) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        std::vector<double> w, x;
        lifted_from_rough(hurst, factors, ratio, w, x);
        std::copy(w.begin(), w.end(), weights_out);
        std::copy(x.begin(), x.end(), rates_out);
    });
}

pe_status pe_rough_plan_compile(const pe_request* carrier, const pe_lifted_heston* model, double maximum_step,
                                const pe_exec_policy* policy, pe_plan** out) {
    if (!carrier || !model || !out) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto plan = std::make_unique<pe_plan>();
    const pe_status st = guarded([&] { compile_rough_plan(plan->host, *carrier, *model, maximum_step, policy); });
    if (st == PE_OK) *out = plan.release();
    return st;
}

pe_status pe_batch_evaluate(const pe_plan* const* plans, uint32_t count, pe_backend backend, pe_result* results) {
    if (!plans || !results || count == 0) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    if (backend != PE_BACKEND_GPU_FP64 && backend != PE_BACKEND_GPU_MIXED) return fail(PE_ERR_INVALID_ARGUMENT, "backend");
    return guarded([&] {
        ensure_init();
        std::vector<HostPlan*> hp;
        for (uint32_t i = 0; i < count; ++i) {
            if (!plans[i]) throw PlanError(PE_ERR_INVALID_ARGUMENT, "null plan");
            hp.push_back(&const_cast<pe_plan*>(plans[i])->host);
        }
        CallStream cs;
        PlanLocks locks(hp);
        // American plans by exercise method, the rest through the BS batch
        std::vector<HostPlan*> groups[3], rest;
        std::vector<size_t> where[3], rest_at;
        for (size_t i = 0; i < hp.size(); ++i) {
            if (hp[i]->rough && hp[i]->american) {
                evaluate_rough_lsm(*hp[i], backend, results + i, nullptr);
            } else if (hp[i]->american) {
                const int g = static_cast<int>(hp[i]->req.lsm.method);
                groups[g].push_back(hp[i]);
                where[g].push_back(i);
            } else {
                rest.push_back(hp[i]);
                rest_at.push_back(i);
            }
        }
        auto scatter = [&](std::vector<HostPlan*>& ps, const std::vector<size_t>& at, auto&& run) {
            if (ps.empty()) return;
            std::vector<pe_result> tmp(ps.size());
            run(ps, tmp.data());
            for (size_t k = 0; k < at.size(); ++k) results[at[k]] = tmp[k];
        };
        for (int g = 0; g < 3; ++g)
            scatter(groups[g], where[g], [&](std::vector<HostPlan*>& ps, pe_result* o) {
                evaluate_lsm(ps, backend, o, nullptr);
            });
        scatter(rest, rest_at, [&](std::vector<HostPlan*>& ps, pe_result* o) { evaluate_bs_batch(ps, backend, o); });
    });
}

pe_status pe_lsm_evaluate(const pe_plan* const* plans, uint32_t count, pe_backend backend, pe_result* results,
                          pe_lsm_report* reports) {
    if (!plans || !results || count == 0) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    if (backend != PE_BACKEND_GPU_FP64 && backend != PE_BACKEND_GPU_MIXED) return fail(PE_ERR_INVALID_ARGUMENT, "backend");
    return guarded([&] {
        ensure_init();
        std::vector<HostPlan*> hp;
        for (uint32_t i = 0; i < count; ++i) {
            if (!plans[i]) throw PlanError(PE_ERR_INVALID_ARGUMENT, "null plan");
            hp.push_back(&const_cast<pe_plan*>(plans[i])->host);
        }
        CallStream cs;
        PlanLocks locks(hp);
        if (hp[0]->rough) {
            // lifted-Heston Bermudans: one launch per plan
            for (size_t i = 0; i < hp.size(); ++i) {
                if (!hp[i]->rough || !hp[i]->american)
                    throw PlanError(PE_ERR_INVALID_ARGUMENT, "mixing lifted-Heston and Black-Scholes American plans");
                evaluate_rough_lsm(*hp[i], backend, results + i, reports ? reports + i : nullptr);
            }
            return;
        }
        for (const HostPlan* q : hp)
            if (q->rough) throw PlanError(PE_ERR_INVALID_ARGUMENT, "mixing lifted-Heston and Black-Scholes American plans");
        evaluate_lsm(hp, backend, results, reports);
    });
}

pe_status pe_lsm_dual_evaluate(const pe_plan* plan, uint32_t continuation_inner_paths,
                               uint32_t exercise_inner_paths, uint64_t inner_seed, pe_backend backend,
                               pe_dual_result* out) {
    if (!plan || !out) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    if (backend != PE_BACKEND_GPU_FP64) return fail(PE_ERR_UNSUPPORTED, "Andersen-Broadie runs in the FP64 backend");
    return guarded([&] {
        ensure_init();
        CallStream cs;
        PlanLocks locks({&const_cast<pe_plan*>(plan)->host});
        evaluate_lsm_dual(&const_cast<pe_plan*>(plan)->host, continuation_inner_paths, exercise_inner_paths,
                          inner_seed, out);
    });
}

pe_status pe_mlp_gradients(const float* params, const float* x, const float* y, uint32_t rows, float* grad_out,
                           double* loss_out) {
    if (!params || !x || !y || !grad_out || rows == 0) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        ensure_init();
        CallStream cs;
        const uint32_t rows_wg = 256;
        const uint32_t wgs = (rows + rows_wg - 1) / rows_wg;
        const size_t np = kMlpParameterCount + 1;
        DeviceBuffer dp, dx, dy, dpart;
        dp.upload(std::vector<float>(params, params + kMlpParameterCount));
        dx.upload(std::vector<float>(x, x + 2 * static_cast<size_t>(rows)));
        dy.upload(std::vector<float>(y, y + rows));
        dpart.reserve(static_cast<size_t>(wgs) * np * sizeof(float));
        check(launch_mlp_gradient(dp.as<float>(), dx.as<float>(), dy.as<float>(), rows, rows_wg, wgs,
                                  dpart.as<float>(), t_stream),
              "MLP gradient launch");
        std::vector<float> part(static_cast<size_t>(wgs) * np);
        check(hipMemcpyAsync(part.data(), dpart.ptr, part.size() * sizeof(float), hipMemcpyDeviceToHost, t_stream),
              "partials copy");
        check(hipStreamSynchronize(t_stream), "hipStreamSynchronize");
        for (size_t k = 0; k < np; ++k) {
            double sum = 0.0;
            for (uint32_t w = 0; w < wgs; ++w) sum += part[static_cast<size_t>(w) * np + k];
            if (k < kMlpParameterCount) grad_out[k] = static_cast<float>(sum);
            else if (loss_out) *loss_out = sum;
        }
    });
}

pe_status pe_mlp_forward(const float* params, const float* x, uint32_t rows, float* out) {
    if (!params || !x || !out || rows == 0) return fail(PE_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        ensure_init();
        CallStream cs;
        DeviceBuffer dp, dx, dout;
        dp.upload(std::vector<float>(params, params + kMlpParameterCount));
        dx.upload(std::vector<float>(x, x + 2 * static_cast<size_t>(rows)));
        dout.reserve(static_cast<size_t>(rows) * sizeof(float));
        check(launch_mlp_forward(dp.as<float>(), dx.as<float>(), rows, dout.as<float>(), t_stream),
              "MLP forward launch");
        check(hipMemcpyAsync(out, dout.ptr, static_cast<size_t>(rows) * sizeof(float), hipMemcpyDeviceToHost,
                             t_stream),
              "output copy");
        check(hipStreamSynchronize(t_stream), "hipStreamSynchronize");
    });
}

pe_status pe_normal_mixed_max_error(uint64_t first_word, uint64_t count, double* max_relative_error,
                                    uint32_t* worst_word) {
    if (!max_relative_error || count == 0 || first_word + count > (1ull << 32))
        return fail(PE_ERR_INVALID_ARGUMENT, "word range");
    return guarded([&] {
        ensure_init();
        CallStream cs;
        constexpr uint32_t kBlocks = 4096;
        DeviceBuffer err, word;
        err.reserve(kBlocks * sizeof(double));
        word.reserve(kBlocks * sizeof(uint32_t));
        check(launch_normal_mixed_error(first_word, count, kBlocks, err.as<double>(), word.as<uint32_t>(), t_stream),
              "accuracy kernel launch");
        std::vector<double> e(kBlocks);
        std::vector<uint32_t> w(kBlocks);
        check(copy_on_stream(e.data(), err.ptr, kBlocks * sizeof(double), hipMemcpyDeviceToHost), "copy");
        check(copy_on_stream(w.data(), word.ptr, kBlocks * sizeof(uint32_t), hipMemcpyDeviceToHost), "copy");
        size_t k = static_cast<size_t>(std::max_element(e.begin(), e.end()) - e.begin());
        *max_relative_error = e[k];
        if (worst_word) *worst_word = w[k];
    });
}

pe_status pe_last_kernel_stats(double* avg_ms, uint32_t* launches) {
    if (avg_ms) *avg_ms = g_kernel_ms;
    if (launches) *launches = g_kernel_launches;
    return PE_OK;
}

}  // extern "C"
