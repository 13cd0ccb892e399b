// Host FP64 plan compilation for the engine.
//
// Mirrors SimulationPlan::compile (crates/pricing/src/engine/compile/simulation.rs:78-640)
// for the single-asset products on the hot path, the local-vol runtime
// (crates/pricing/src/engine/compile/local_vol.rs:14-103, processes/local_vol.rs:166-237),
// the RQMC scramble table (crates/pricing/src/engine/sampling/qmc.rs:190-369) and the
// Brownian-bridge instruction order (crates/pricing/src/engine/sampling/bridge.rs:164-467).
// All arithmetic is FP64 with the reference's expressions; the device tables
// carry the results.
#include "pe_plan.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <queue>

#include "pe_device.hpp"

namespace pe {

namespace {

constexpr double kRelativeSpotBump = 1.0e-4;  // crates/pricing/src/api/mc_result.rs:288
constexpr double kVolatilityBump = 1.0e-4;    // crates/pricing/src/api/mc_result.rs:289

[[noreturn]] void invalid(const std::string& what) { throw PlanError(PE_ERR_INVALID_ARGUMENT, what); }
[[noreturn]] void unsupported(const std::string& what) { throw PlanError(PE_ERR_UNSUPPORTED, what); }

// Index of the last node ordered at or before `value`, clamped to the last cell.
size_t lower_cell(const std::vector<double>& nodes, double value) {
    const size_t count = static_cast<size_t>(std::upper_bound(nodes.begin(), nodes.end(), value) - nodes.begin());
    const size_t idx = count == 0 ? 0 : count - 1;
    return std::min(idx, nodes.size() - 2);
}

void check_sorted(const std::vector<double>& v, bool reject_negative, const char* what) {
    if (v.size() < 2) invalid(std::string(what) + ": needs at least two nodes");
    for (size_t i = 0; i < v.size(); ++i) {
        if (!std::isfinite(v[i]) || (reject_negative && v[i] < 0.0)) invalid(std::string(what) + ": invalid node");
        if (i > 0 && !(v[i] > v[i - 1])) invalid(std::string(what) + ": nodes must increase");
    }
}

// Sorted unique fixing days on or after valuation (CompiledPayoff::terminal_observations,
// crates/pricing/src/engine/payoff/tape.rs:229-247).
std::vector<int32_t> fixing_days(const pe_product& p) {
    std::vector<int32_t> d;
    auto add = [&](int32_t day) { if (day >= 0) d.push_back(day); };
    switch (p.kind) {
        case PE_PRODUCT_EUROPEAN:
        case PE_PRODUCT_DIGITAL:
            add(p.expiry_day);
            break;
        case PE_PRODUCT_ASIAN:
            for (uint32_t i = 0; i < p.date_count; ++i)
                if (!(p.fixings && !std::isnan(p.fixings[i]))) add(p.dates[i]);
            break;
        case PE_PRODUCT_BARRIER:
            add(p.expiry_day);
            if (!(p.continuous && p.historical_hit == 1))
                for (uint32_t i = 0; i < p.date_count; ++i) add(p.dates[i]);
            break;
        case PE_PRODUCT_LOOKBACK:
        case PE_PRODUCT_AMERICAN:
            for (uint32_t i = 0; i < p.date_count; ++i) add(p.dates[i]);
            break;
        default:
            unsupported("product kind");
    }
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    return d;
}

}  // namespace

void HostCurve::init(const pe_curve& c) {
    if (c.count < 2 || !c.times || !c.discount_factors) invalid("curve needs two pillars");
    times.assign(c.times, c.times + c.count);
    if (times[0] != 0.0) invalid("curve must start at t = 0");
    if (c.discount_factors[0] != 1.0) invalid("curve must start at DF = 1");
    log_dfs.resize(c.count);
    for (uint32_t i = 0; i < c.count; ++i) {
        if (!std::isfinite(times[i]) || times[i] < 0.0) invalid("curve pillar time");
        if (!std::isfinite(c.discount_factors[i]) || c.discount_factors[i] <= 0.0) invalid("discount factor");
        if (i > 0 && !(times[i] > times[i - 1])) invalid("curve pillars must increase");
        log_dfs[i] = std::log(c.discount_factors[i]);
    }
    times[0] = 0.0;
}

// CurveRegion of a query time (LogLinearDiscountCurve::evaluate, curve.rs:165-190).
pe_curve_region HostCurve::region(double t) const {
    if (t == 0.0) t = 0.0;
    const auto it = std::lower_bound(times.begin(), times.end(), t);
    if (it != times.end() && *it == t) return PE_CURVE_PILLAR;
    return it != times.end() ? PE_CURVE_INTERPOLATED : PE_CURVE_RIGHT_EXTRAPOLATED;
}

double HostCurve::log_discount(double t) const {
    if (!std::isfinite(t) || t < 0.0) invalid("curve query time");
    if (t == 0.0) t = 0.0;
    const auto it = std::lower_bound(times.begin(), times.end(), t);
    const size_t right = static_cast<size_t>(it - times.begin());
    auto segment = [&](size_t l, size_t r) {
        const double w = (t - times[l]) / (times[r] - times[l]);
        return log_dfs[l] + w * (log_dfs[r] - log_dfs[l]);
    };
    if (it != times.end() && *it == t) return log_dfs[right];
    if (right < times.size()) return segment(right - 1, right);
    return segment(times.size() - 2, times.size() - 1);
}

// EquityForward::evaluate (crates/pricing/src/market/forward.rs:129-161), no discrete dividends.
double HostPlan::forward_at_spot(double t, double s) const {
    const double f = s * std::exp(dividend_curve.log_discount(t) - discount_curve.log_discount(t));
    if (!std::isfinite(f) || f <= 0.0) invalid("non-finite forward");
    return f;
}

// ------------------------------------------------------------ dividends --
// AffineDividendTransform::new_with_carry_curves + compile_reserve
// (crates/pricing/src/market/dividend.rs:318-361, 622-668).
void HostDividends::compile(double s0, const pe_dividend_quote* q, uint32_t n, const HostCurve* p,
                            const HostCurve* d) {
    spot = s0;
    discount = p;
    dividend = d;
    events.clear();
    for (uint32_t i = 0; i < n; ++i) {
        const double t = q[i].ex_time, cash = q[i].fixed_cash, beta = q[i].beta;
        if (!std::isfinite(t) || t < 0.0) invalid("dividend ex-time must be finite and non-negative");
        if (!(std::isfinite(cash) && cash >= 0.0)) invalid("dividend cash must be finite and non-negative");
        if (!(std::isfinite(beta) && beta >= 0.0 && beta < 1.0)) invalid("dividend proportion must be in [0, 1)");
        if (i > 0 && !(t > q[i - 1].ex_time)) invalid("dividend ex-times must increase strictly");
        const double alpha = cash / s0;
        if (!std::isfinite(alpha)) invalid("non-finite dividend alpha");
        events.push_back({t, cash, alpha, beta});
    }
    double proportional = 1.0, r = 0.0;
    for (const Event& e : events) {
        proportional *= 1.0 - e.beta;
        const double ratio = std::exp(p->log_discount(e.ex_time)) / std::exp(d->log_discount(e.ex_time));
        r += e.cash / proportional * ratio;
    }
    const double residual = spot - r;
    if (!std::isfinite(r) || !std::isfinite(residual) || residual <= 0.0) invalid("non-positive escrowed spot");
    reserve = r;
    // event_timeline (dividend.rs:524-547)
    timeline.clear();
    HostCoord c = initial();
    double prev = 0.0;
    long last_cash = -1;
    for (size_t i = 0; i < events.size(); ++i)
        if (events[i].cash != 0.0) last_cash = static_cast<long>(i);
    for (size_t i = 0; i < events.size(); ++i) {
        const HostCoord before = carry(c, prev, events[i].ex_time);
        HostCoord after{(1.0 - events[i].beta) * before.a - events[i].alpha, (1.0 - events[i].beta) * before.b,
                        (1.0 - events[i].beta) * before.ss};
        if (!std::isfinite(after.a) || !std::isfinite(after.b)) invalid("non-finite dividend transform");
        if (last_cash < 0 || static_cast<long>(i) >= last_cash) after.a = 0.0;
        timeline.push_back({events[i].ex_time, before, after});
        c = after;
        prev = events[i].ex_time;
    }
}

HostCoord HostDividends::initial() const { return HostCoord{reserve / spot, (spot - reserve) / spot, 1.0}; }

HostCoord HostDividends::carry(HostCoord c, double start, double end) const {
    if (c.a == 0.0 || start == end) return c;
    const double g = dividend->log_discount(end) - discount->log_discount(end) - dividend->log_discount(start) +
                     discount->log_discount(start);
    c.a *= std::exp(g);
    if (!std::isfinite(c.a)) invalid("non-finite carried affine coordinate");
    return c;
}

// coordinate_after_time (dividend.rs:470-497)
HostCoord HostDividends::after_time(double t) const {
    HostCoord c = initial();
    double prev = 0.0;
    for (const Event& e : events) {
        if (!(e.ex_time <= t)) break;
        c = carry(c, prev, e.ex_time);
        c = HostCoord{(1.0 - e.beta) * c.a - e.alpha, (1.0 - e.beta) * c.b, (1.0 - e.beta) * c.ss};
        if (!std::isfinite(c.a) || !std::isfinite(c.b)) invalid("non-finite dividend transform");
        prev = e.ex_time;
    }
    c = carry(c, prev, t);
    bool all = true;
    for (const Event& e : events)
        if (!(e.ex_time <= t || e.cash == 0.0)) all = false;
    if (all) c.a = 0.0;
    return c;
}

// ----------------------------------------------------------- bridge order --
void compile_bridge_program(const std::vector<double>& t, std::vector<BridgeOp>& program,
                            std::vector<double>& inv_sqrt_dt, uint32_t& depth) {
    if (t.size() < 2) invalid("Brownian bridge needs two nodes");
    if (t[0] != 0.0 || std::signbit(t[0])) invalid("Brownian bridge must start at +0");
    for (size_t i = 0; i + 1 < t.size(); ++i)
        if (!(t[i + 1] > t[i]) || !std::isfinite(t[i + 1])) invalid("Brownian bridge times must increase");
    inv_sqrt_dt.resize(t.size() - 1);
    for (size_t i = 0; i + 1 < t.size(); ++i) inv_sqrt_dt[i] = 1.0 / std::sqrt(t[i + 1] - t[i]);

    struct Cand {
        double variance, time;
        size_t node, left, right;
    };
    // Ord of IntervalCandidate (bridge.rs:137-146): variance, then earlier time, then lower node.
    auto less = [](const Cand& a, const Cand& b) {
        if (a.variance != b.variance) return a.variance < b.variance;
        if (a.time != b.time) return a.time > b.time;
        return a.node > b.node;
    };
    auto candidate = [&](size_t l, size_t r, Cand* out) {
        if (r <= l + 1) return false;
        const double den = t[r] - t[l];
        bool have = false;
        Cand best{};
        for (size_t m = l + 1; m < r; ++m) {
            const double var = ((t[m] - t[l]) * (t[r] - t[m])) / den;
            if (!std::isfinite(var) || var <= 0.0) invalid("Brownian bridge conditional variance");
            const Cand c{var, t[m], m, l, r};
            if (!have || !less(c, best)) { best = c; have = true; }
        }
        *out = best;
        return true;
    };
    struct Split {
        size_t node;
        uint32_t rank;
        double lw, rw, sd;
    };
    std::map<std::pair<size_t, size_t>, Split> splits;
    const size_t n = t.size() - 1;
    std::priority_queue<Cand, std::vector<Cand>, decltype(less)> heap(less);
    Cand c;
    if (candidate(0, n, &c)) heap.push(c);
    uint32_t rank = 1;
    while (!heap.empty()) {
        c = heap.top();
        heap.pop();
        const double den = t[c.right] - t[c.left];
        splits[{c.left, c.right}] = Split{c.node, rank++, (t[c.right] - t[c.node]) / den,
                                          (t[c.node] - t[c.left]) / den, std::sqrt(c.variance)};
        Cand a, b;
        if (candidate(c.left, c.node, &a)) heap.push(a);
        if (candidate(c.node, c.right, &b)) heap.push(b);
    }
    // In-order traversal of the interval tree: values come out in time order.
    program.clear();
    auto op = [](int32_t push, int32_t rank, int32_t terminal, int32_t step, double lw, double rw, double sd) {
        return BridgeOp{push, rank, terminal, step, lw, rw, sd, static_cast<float>(lw), static_cast<float>(rw),
                        static_cast<float>(sd), 0.0f};
    };
    program.push_back(op(1, 0, 1, 0, 0.0, 0.0, std::sqrt(t[n])));
    std::vector<size_t> stack{n};
    depth = 1;
    size_t left = 0;
    while (!stack.empty()) {
        const size_t right = stack.back();
        if (right == left + 1) {
            program.push_back(op(0, 0, 0, static_cast<int32_t>(left), 0.0, 0.0, 0.0));
            stack.pop_back();
            left = right;
        } else {
            const auto it = splits.find({left, right});
            if (it == splits.end()) invalid("Brownian bridge interval without a split");
            const Split& s = it->second;
            program.push_back(op(1, static_cast<int32_t>(s.rank), 0, 0, s.lw, s.rw, s.sd));
            stack.push_back(s.node);
            depth = std::max<uint32_t>(depth, static_cast<uint32_t>(stack.size()));
        }
    }
}

// ------------------------------------------------------------ Sobol table --
// [scramble][dim][33]: LMS-scrambled direction words (MSB-first unit lower
// triangular rows from Philox domain RqmcScramble) and the digital shift.
extern "C" const unsigned char pe_jk_directions_begin[];
extern "C" const unsigned char pe_jk_directions_end[];

// compile_scramble (qmc.rs:107-133): digital shift and MSB-first unit lower
// triangular rows of one (scramble, dimension) from Philox domain RqmcScramble.
static uint32_t scramble_rows(uint64_t seed, uint32_t s, uint32_t d, uint32_t (&rows)[32]) {
    auto word = [seed](uint64_t scramble, uint32_t dim) {
        const U4 b = philox_block(seed, scramble, dim >> 2, 2u);
        return lane_of(b, dim & 3u);
    };
    const uint32_t block = d * 32u;
    const uint32_t shift = word(s, block);
    rows[0] = 0x80000000u;
    for (uint32_t r = 1; r < 32; ++r) {
        const uint32_t raw = word(s, block + r);
        rows[r] = (raw & (0xffffffffu << (32 - r))) | (1u << (31 - r));
    }
    return shift;
}

// RqmcPlan checksums (qmc.rs:236-259): BLAKE3 of the direction asset, and of
// every scramble's shift and rows as big-endian words, scramble-major.
void rqmc_checksums(uint64_t seed, uint32_t scrambles, uint32_t dims, uint8_t (&direction)[32],
                    uint8_t (&scramble)[32]) {
    static const std::vector<uint8_t> asset_hash = [] {
        uint8_t h[32];
        Blake3 b;
        b.update(pe_jk_directions_begin, static_cast<size_t>(pe_jk_directions_end - pe_jk_directions_begin));
        b.finalize(h);
        return std::vector<uint8_t>(h, h + 32);
    }();
    std::copy(asset_hash.begin(), asset_hash.end(), direction);
    Blake3 b;
    auto be = [&](uint32_t v) {
        const uint8_t bytes[4] = {static_cast<uint8_t>(v >> 24), static_cast<uint8_t>(v >> 16),
                                  static_cast<uint8_t>(v >> 8), static_cast<uint8_t>(v)};
        b.update(bytes, 4);
    };
    for (uint32_t s = 0; s < scrambles; ++s)
        for (uint32_t d = 0; d < dims; ++d) {
            uint32_t rows[32];
            be(scramble_rows(seed, s, d, rows));
            for (uint32_t r : rows) be(r);
        }
    b.finalize(scramble);
}

std::vector<uint32_t> build_sobol_table(uint64_t seed, uint32_t scrambles, uint32_t dims) {
    uint32_t available = 0;
    const uint32_t* words = direction_words(&available);
    if (dims == 0) invalid("zero Sobol dimension");
    if (dims > available) throw PlanError(PE_ERR_SIZE_LIMIT, "Sobol dimension exceeds the Joe--Kuo limit");
    std::vector<uint32_t> table(static_cast<size_t>(scrambles) * dims * 33u);
    for (uint32_t s = 0; s < scrambles; ++s) {
        for (uint32_t d = 0; d < dims; ++d) {
            uint32_t rows[32];
            const uint32_t shift = scramble_rows(seed, s, d, rows);
            uint32_t* out = &table[(static_cast<size_t>(s) * dims + d) * 33u];
            const uint32_t* dir = words + static_cast<size_t>(d) * 32u;
            for (int k = 0; k < 32; ++k) {
                uint32_t v = 0;
                for (int r = 0; r < 32; ++r) v |= (static_cast<uint32_t>(__builtin_popcount(dir[k] & rows[r])) & 1u) << (31 - r);
                out[k] = v;
            }
            out[32] = shift;
        }
    }
    return table;
}

// ------------------------------------------------------------ BS trade --
// Per-observation affine coordinates, pre-dividend coordinates on ex-dates
// that the payoff monitors (compile/simulation.rs:115-130, 308-338), the
// continuous-barrier runtime (:339-414) and the kernel's trade descriptor.
void compile_bs_trade(HostPlan& p, double monitoring_end) {
    const pe_product& pr = p.req.product;
    const size_t n = p.obs_times.size();
    p.obs_coord.assign(n, HostCoord{});
    p.obs_pre.assign(n, HostCoord{});
    p.obs_flags.assign(n, 0u);
    bool has_div = false;
    for (size_t i = 0; i < n; ++i) {
        const double t = p.obs_times[i];
        if (!p.dividends.empty()) p.obs_coord[i] = p.dividends.after_time(t);
        const HostCoord& c = p.obs_coord[i];
        const double spot_contract = c.a * p.spot + c.b * p.obs_forwards[i];
        if (!std::isfinite(spot_contract) || spot_contract <= 0.0) invalid("non-finite dividend forward");
        if (c.a != 0.0 || c.b != 1.0 || c.ss != 1.0) has_div = true;
        const HostDividends::Entry* entry = nullptr;
        for (const auto& e : p.dividends.timeline)
            if (e.ex_time == t) entry = &e;
        bool jump = false;
        if (entry && pr.kind == PE_PRODUCT_BARRIER && !p.continuous && p.obs_days[i] != kNoDay)
            for (uint32_t k = 0; k < pr.date_count; ++k)
                if (pr.dates[k] == p.obs_days[i]) jump = true;
        const bool graph_pre = p.graph && p.obs_days[i] != kNoDay &&
                               std::binary_search(p.graph_tape.pre_days.begin(), p.graph_tape.pre_days.end(),
                                                  p.obs_days[i]);
        if (graph_pre && !entry) unsupported("pre-dividend graph observation on a date without a dividend");
        if (entry && (p.continuous || jump || graph_pre)) {
            p.obs_flags[i] |= kObsPre;
            p.obs_pre[i] = entry->before;
            has_div = true;
        }
        if (jump) p.obs_flags[i] |= kObsJump;
        if (p.obs_weight.size() == n && p.obs_weight[i] != 0.0) p.obs_flags[i] |= kObsFixing;
        if (p.continuous && t <= monitoring_end) p.obs_flags[i] |= kObsBridge;
    }
    p.single_terminal = n == 1 && p.obs_days[0] == pr.expiry_day && !(p.obs_flags[0] & kObsPre);
    p.resolved = false;
    p.resolved_survival = 1.0;
    if (p.continuous) {
        if (pr.historical_hit == 1) { p.resolved = true; p.resolved_survival = 0.0; }
        else if (monitoring_end < 0.0) { p.resolved = true; p.resolved_survival = 1.0; }
        if (p.product.smoothing && !pr.direction_down && pr.historical_hit != 1 && p.dates.back() >= 0) {
            std::vector<HostCoord> coords{HostCoord{}};
            coords.insert(coords.end(), p.obs_coord.begin(), p.obs_coord.end());
            for (size_t i = 0; i < n; ++i)
                if (p.obs_flags[i] & kObsPre) coords.push_back(p.obs_pre[i]);
            for (const HostCoord& c : coords)
                if (p.product.half_width >= pr.barrier - c.a * p.spot)
                    invalid("smoothing half width reaches the transformed spot barrier");
        }
    }
    p.bs_obs.assign(n, BsObs{});
    double prev = 0.0;
    for (size_t i = 0; i < n; ++i) {
        BsObs& o = p.bs_obs[i];
        o.time = p.obs_times[i];
        o.sqrt_step = std::sqrt(std::max(o.time - prev, 0.0));
        prev = o.time;
        o.forward = p.obs_forwards[i];
        o.weight = p.obs_weight.size() == n ? p.obs_weight[i] : 0.0;
        o.a = p.obs_coord[i].a; o.b = p.obs_coord[i].b; o.ss = p.obs_coord[i].ss;
        o.pa = p.obs_pre[i].a; o.pb = p.obs_pre[i].b; o.pss = p.obs_pre[i].ss;
        o.flags = p.obs_flags[i];
    }
    BsTrade& t = p.trade;
    std::memset(&t, 0, sizeof(t));
    t.product = p.product;
    const pe_engine& e = p.req.engine;
    t.seed = e.seed;
    t.units = e.units;
    t.s0 = p.spot;
    t.reserve = p.dividends.reserve;
    const HostCoord c0 = p.dividends.empty() ? HostCoord{} : p.dividends.initial();
    t.init_a = c0.a;
    t.init_b = c0.b;
    t.gamma_bump = p.gamma_bump;
    t.spot_bump = p.spot_bump;
    t.vol_bump = p.vol_bump;
    t.resolved_survival = p.resolved_survival;
    t.n_obs = static_cast<uint32_t>(n);
    t.sobol_dims = p.sobol_dims;
    t.rqmc = e.kind == PE_ENGINE_RQMC ? 1u : 0u;
    t.antithetic = e.antithetic ? 1u : 0u;
    t.single_terminal = p.single_terminal ? 1u : 0u;
    t.request_gamma = p.request_gamma ? 1u : 0u;
    t.zero_variance = p.total_variance == 0.0 ? 1u : 0u;
    t.has_dividends = has_div || !p.dividends.empty() ? 1u : 0u;
    t.continuous = p.continuous ? 1u : 0u;
    t.resolved = p.resolved ? 1u : 0u;
    t.expiry_obs = static_cast<uint32_t>(p.expiry_obs);
    t.risk = p.risk ? 1u : 0u;
    t.graph_output = p.graph_tape.output;
    t.graph_regs = p.graph_tape.regs;
    auto variant = [&](int k, double spot, double vol) {
        t.variants[k] = Variant{spot, vol, spot / p.spot, static_cast<float>(spot), static_cast<float>(vol),
                                static_cast<float>(vol * 1.4426950408889634), static_cast<float>(1.0 / spot)};
    };
    variant(0, p.spot, p.vol);
    uint32_t nv = 1;
    if (p.risk) {
        variant(1, p.spot - p.spot_bump, p.vol);
        variant(2, p.spot + p.spot_bump, p.vol);
        nv = 3;
        if (p.vol_bump != 0.0) {
            variant(3, p.spot, p.vol - p.vol_bump);
            variant(4, p.spot, p.vol + p.vol_bump);
            nv = 5;
        }
    }
    t.n_variants = nv;
    // per-observation coordinates at the spot-bumped variants, with the
    // kernel's at_spot arithmetic (cash held fixed: dividend.rs:163-173)
    for (BsObs& o : p.bs_obs) {
        for (int k = 0; k < 2; ++k) {
            const double spot = nv >= 3 ? t.variants[k + 1].spot : p.spot;
            const double original = p.spot;
            if (original == spot) {
                o.va[k] = o.a; o.vb[k] = o.b; o.vpa[k] = o.pa; o.vpb[k] = o.pb;
            } else {
                o.va[k] = o.a * original / spot;
                o.vb[k] = o.ss - (o.ss - o.b) * original / spot;
                o.vpa[k] = o.pa * original / spot;
                o.vpb[k] = o.pss - (o.pss - o.pb) * original / spot;
            }
        }
    }
    // FP32 constants of the mixed kernel, from the FP64 values above
    const double log2e = 1.4426950408889634;
    const double spots[3] = {p.spot, nv >= 3 ? t.variants[1].spot : p.spot, nv >= 3 ? t.variants[2].spot : p.spot};
    const double vols[3] = {p.vol, nv >= 5 ? t.variants[3].vol : p.vol, nv >= 5 ? t.variants[4].vol : p.vol};
    const double barrier = p.product.barrier;
    auto f32 = [](double x) { return static_cast<float>(x); };
    p.bs_obs32.assign(p.bs_obs.size(), BsObs32{});
    double prev_bridge = 0.0;   // time of the previous bridge observation (valuation: 0)
    for (size_t i = 0; i < p.bs_obs.size(); ++i) {
        const BsObs& o = p.bs_obs[i];
        BsObs32& q = p.bs_obs32[i];
        q.flags = o.flags;
        q.time = f32(o.time);
        q.sqrt_step = f32(o.sqrt_step);
        q.weight = f32(o.weight);
        q.sqrt_time = f32(std::sqrt(o.time));
        q.ss = f32(o.ss);
        q.pss = f32(o.pss);
        for (int si = 0; si < 3; ++si) {
            const double spot = spots[si];
            const double A = si ? o.va[si - 1] : o.a, B = si ? o.vb[si - 1] : o.b;
            const double PA = si ? o.vpa[si - 1] : o.pa, PB = si ? o.vpb[si - 1] : o.pb;
            q.fwd[si] = f32(o.forward * (spot / p.spot));
            q.cash[si] = f32(A * spot);
            q.b[si] = f32(B);
            q.pcash[si] = f32(PA * spot);
            q.pb[si] = f32(PB);
            q.ssos[si] = f32(o.ss / spot);
            q.pssos[si] = f32(o.pss / spot);
            q.a[si] = f32(A);
            q.pa[si] = f32(PA);
            const double bar = B != 0.0 ? (barrier - A * spot) / B : 0.0;
            const double pbar = PB != 0.0 ? (barrier - PA * spot) / PB : 0.0;
            q.bar[si] = f32(bar);
            q.pbar[si] = f32(pbar);
            q.itsb[si] = B * bar != 0.0 ? f32(1.0 / (B * bar)) : 0.0f;
            q.pitsb[si] = PB * pbar != 0.0 ? f32(1.0 / (PB * pbar)) : 0.0f;
        }
        const bool bridge = (o.flags & kObsBridge) != 0u;
        const double dt = bridge ? o.time - prev_bridge : 0.0;
        if (bridge) prev_bridge = o.time;
        for (int vi = 0; vi < 3; ++vi) {
            const double vol = vols[vi];
            q.mh2[vi] = f32(-0.5 * (vol * vol * o.time) * log2e);
            q.nvt[vi] = f32(-vol * o.time);
            q.sd2[vi] = f32(vol * std::sqrt(o.time) * log2e);
            const double var = vol * vol;
            const double iv = (0.5 * var + 0.5 * var) * dt;
            q.iv[vi] = f32(iv);
            q.liv[vi] = iv > 0.0 ? f32(std::log(iv)) : 0.0f;
            q.iiv[vi] = iv > 0.0 ? f32(1.0 / iv) : 0.0f;
        }
    }
}

// ------------------------------------------------------------ time grid --
std::vector<double> compile_time_grid(std::vector<double> events, double maximum_step) {
    if (!std::isfinite(maximum_step) || maximum_step <= 0.0) invalid("maximum step must be positive");
    std::sort(events.begin(), events.end());
    events.erase(std::unique(events.begin(), events.end()), events.end());
    if (events.empty() || events[0] != 0.0) events.insert(events.begin(), 0.0);
    std::vector<double> nodes(1, 0.0);
    for (size_t i = 0; i + 1 < events.size(); ++i) {
        const double l = events[i], rr = events[i + 1];
        const double interval = rr - l;
        const double raw = std::ceil(interval / maximum_step);
        if (!std::isfinite(raw) || raw > 1e9) throw PlanError(PE_ERR_SIZE_LIMIT, "time grid step count");
        const size_t steps = std::max<size_t>(1, static_cast<size_t>(raw));
        const double sub = interval / static_cast<double>(steps);
        for (size_t k = 1; k < steps; ++k) nodes.push_back(l + sub * static_cast<double>(k));
        nodes.push_back(rr);
    }
    return nodes;
}

// ------------------------------------------------------- lifted Heston --
// gamma_half_to_two (crates/pricing-numerics/src/fractional.rs:1-30): Lanczos
// with g = 7 and Godfrey's nine published coefficients, on [1/2, 2].
static double lanczos_gamma(double x) {
    if (!std::isfinite(x) || !(x >= 0.5 && x <= 2.0)) invalid("gamma argument outside [1/2, 2]");
    if (x == 1.0 || x == 2.0) return 1.0;
    static const double c[8] = {676.5203681218851,     -1259.1392167224028,   771.3234287776531,
                                -176.6150291621406,    12.507343278686905,    -0.13857109526572012,
                                9.984369578019572e-6, 1.5056327351493116e-7};
    const double z = x - 1.0;
    double sum = 0.9999999999998099;
    for (int i = 0; i < 8; ++i) sum += c[i] / (z + static_cast<double>(i) + 1.0);
    const double t = z + 7.5;
    return std::sqrt(2.0 * M_PI) * std::pow(t, z + 0.5) * std::exp(-t) * sum;
}

// LiftedHeston::from_rough (crates/pricing/src/models/rough_volatility.rs:564-606).
void lifted_from_rough(double hurst, uint32_t factors, double ratio, std::vector<double>& weights,
                       std::vector<double>& rates) {
    if (factors == 0 || factors > 256 || !std::isfinite(ratio) || ratio <= 1.0) invalid("lifted Heston quadrature");
    if (!std::isfinite(hurst) || hurst <= 0.0 || hurst > 0.5) invalid("Hurst index must be in (0, 1/2]");
    weights.clear();
    rates.clear();
    if (hurst == 0.5) {
        weights.push_back(1.0);
        rates.push_back(0.0);
        return;
    }
    const double alpha = hurst + 0.5;
    const double p = 1.0 - alpha;
    const double log_ratio = std::log(ratio);
    const double denominator = lanczos_gamma(alpha) * lanczos_gamma(2.0 - alpha);
    const double mass = std::expm1(p * log_ratio) / denominator;
    const double rate = p / (1.0 + p) * std::expm1((1.0 + p) * log_ratio) / std::expm1(p * log_ratio);
    for (uint32_t j = 0; j < factors; ++j) {
        const double log_lower = (static_cast<double>(j) - static_cast<double>(factors) / 2.0) * log_ratio;
        weights.push_back(mass * std::exp(p * log_lower));
        rates.push_back(rate * std::exp(log_lower));
    }
}

// RoughVolatilityPricingPlan::compile (crates/pricing/src/engine/risk/rough_volatility.rs:84-190).
void compile_rough_plan(HostPlan& p, const pe_request& carrier, const pe_lifted_heston& m, double maximum_step,
                        const pe_exec_policy* policy) {
    if (carrier.model.kind != PE_MODEL_BLACK_SCHOLES) invalid("rough extension requires a Black-Scholes request carrier");
    if (carrier.risk.delta || carrier.risk.vega || carrier.risk.gamma) unsupported("rough pricing is price-only");
    if (carrier.market.dividend_count) unsupported("discrete dividends under lifted Heston");
    if (carrier.product.kind == PE_PRODUCT_BARRIER && carrier.product.continuous)
        unsupported("continuous barriers in the rough adapter");
    const bool american = carrier.product.kind == PE_PRODUCT_AMERICAN;
    if (carrier.product.kind != PE_PRODUCT_EUROPEAN && carrier.product.kind != PE_PRODUCT_DIGITAL && !american)
        unsupported("lifted Heston on the device prices European, Digital and Bermudan (American) payoffs");
    if (american) {
        const pe_lsm& l = carrier.lsm;
        if (l.method != PE_LSM_POLY_CPQR && l.method != PE_LSM_MLP)
            unsupported("lifted Heston early exercise: the CPQR monomial or the network policy");
        if (l.max_degree < 1 || l.max_degree > 2) unsupported("lifted Heston early exercise: total degree 1 or 2");
        if (carrier.engine.brownian_bridge || l.training_engine.brownian_bridge)
            invalid("no Brownian bridge for lifted Heston early exercise");
    }
    compile_host_plan(p, carrier, policy);
    HostLifted& h = p.lifted;
    h.v0 = m.initial_variance;
    h.kappa = m.mean_reversion;
    h.theta = m.long_run_variance;
    h.nu = m.vol_of_vol;
    h.rho = m.correlation;
    for (double v : {h.v0, h.kappa, h.theta, h.nu})
        if (!std::isfinite(v) || v < 0.0) invalid("lifted Heston parameters must be finite and non-negative");
    if (!std::isfinite(h.rho) || h.rho < -1.0 || h.rho > 1.0) invalid("lifted Heston correlation");
    if (m.factor_count == 0 || m.factor_count > 256 || !m.weights || !m.rates) invalid("lifted Heston factor count");
    if (m.factor_count > static_cast<uint32_t>(kMaxRoughFactors))
        throw PlanError(PE_ERR_SIZE_LIMIT, "the device keeps at most 32 lifted Heston factors");
    h.weights.assign(m.weights, m.weights + m.factor_count);
    h.rates.assign(m.rates, m.rates + m.factor_count);
    for (size_t i = 0; i < h.weights.size(); ++i) {
        if (!std::isfinite(h.weights[i]) || h.weights[i] <= 0.0) invalid("lifted Heston weight");
        if (!std::isfinite(h.rates[i]) || h.rates[i] < 0.0) invalid("lifted Heston rate");
    }
    if (!std::isfinite(p.expiry) || p.expiry <= 0.0 || !std::isfinite(maximum_step) || maximum_step <= 0.0)
        invalid("rough positive horizon and step");
    p.rough = true;
    p.rough_max_step = maximum_step;
    p.rough_events = p.obs_times;
    p.rough_events.push_back(0.0);
    p.rough_events.push_back(p.expiry);
    std::sort(p.rough_events.begin(), p.rough_events.end());
    p.rough_events.erase(std::unique(p.rough_events.begin(), p.rough_events.end()), p.rough_events.end());
    double estimated = 0.0;
    for (size_t i = 0; i + 1 < p.rough_events.size(); ++i)
        estimated += std::ceil((p.rough_events[i + 1] - p.rough_events[i]) / maximum_step);
    if (!std::isfinite(estimated) || estimated > 65536.0) throw PlanError(PE_ERR_SIZE_LIMIT, "rough time grid resource limit");
    // RQMC dimensions of the plan's own grid: two normals per step
    const size_t steps = compile_time_grid(p.rough_events, maximum_step).size() - 1;
    compile_diagnostics(p, policy, static_cast<uint32_t>(2 * steps));
    if (american) {
        const pe_engine& te = p.req.lsm.training_engine;
        p.train_diag = pe_mc_diagnostics{};
        if (te.kind == PE_ENGINE_RQMC) {
            p.train_diag.has_checksums = 1;
            rqmc_checksums(te.seed, te.scrambles, static_cast<uint32_t>(2 * steps), p.train_diag.direction_checksum,
                           p.train_diag.scramble_checksum);
        }
    }
}

// ------------------------------------------------------------------ plan --
// Early-exercise runtime (execute_early_exercise, crates/pricing/src/engine/risk/early_exercise.rs:24-300):
// per exercise date the exact-step constants and DF(date); the training and
// valuation engines and the regression settings of the LsmConfig.
void compile_lsm_trade(HostPlan& p) {
    const pe_request& r = p.req;
    const pe_lsm& l = r.lsm;
    if (p.lv) unsupported("early exercise under local volatility on the device");
    if (l.method != PE_LSM_POLY_CPQR && l.method != PE_LSM_POLY_LAGUERRE && l.method != PE_LSM_MLP)
        invalid("LSM method");
    if (l.max_degree + 1 > static_cast<uint32_t>(kLsmMaxBasis)) unsupported("LSM basis degree above 3 on the device");
    const pe_engine& te = l.training_engine;
    const pe_engine& ve = r.engine;
    if (te.kind != ve.kind) invalid("training and valuation engines must be of one kind");
    if (te.units == 0) invalid("zero training units");
    if (te.kind == PE_ENGINE_RQMC && (te.scrambles == 0 || te.units > 0xffffffffull))
        invalid("RQMC needs scrambles and at most 2^32 points");
    if (!std::isfinite(l.abs_rank_tolerance) || l.abs_rank_tolerance < 0.0 || !std::isfinite(l.rel_rank_tolerance) ||
        l.rel_rank_tolerance < 0.0 || !std::isfinite(l.itm_abs_tolerance) || l.itm_abs_tolerance < 0.0)
        invalid("LSM tolerances must be finite and non-negative");
    const size_t n = p.obs_times.size();
    if (n == 0 || n > static_cast<size_t>(kLsmMaxDates)) throw PlanError(PE_ERR_SIZE_LIMIT, "exercise date count");
    const uint64_t train_mult = te.antithetic ? 2 : 1;
    const uint64_t train_paths = te.units * (te.kind == PE_ENGINE_RQMC ? te.scrambles : 1u) * train_mult;
    if (l.max_matrix_elements && train_paths > l.max_matrix_elements)
        invalid("regression matrix element limit");
    p.lsm_dates.assign(n, LsmDate{});
    for (size_t i = 0; i < n; ++i) {
        LsmDate& d = p.lsm_dates[i];
        const BsObs& o = p.bs_obs[i];
        d.time = o.time;
        d.sqrt_step = o.sqrt_step;
        d.forward = o.forward;
        d.discount = std::exp(p.discount_curve.log_discount(static_cast<double>(p.obs_days[i]) / 365.0));
        d.a = o.a;
        d.b = o.b;
        d.ss = o.ss;
        d.m_half_tv = -0.5 * (p.vol * p.vol * o.time);
        d.inv_discount = 1.0 / d.discount;
        d.f_sqrt_step = static_cast<float>(d.sqrt_step);
        d.f_forward = static_cast<float>(d.forward);
        d.f_m_half_tv = static_cast<float>(d.m_half_tv);
    }
    LsmTrade& t = p.lsm_trade;
    std::memset(&t, 0, sizeof(t));
    auto engine = [&](const pe_engine& e, uint32_t domain) {
        LsmEngine g{};
        g.seed = e.seed;
        g.units = e.units;
        g.scrambles = e.kind == PE_ENGINE_RQMC ? e.scrambles : 1u;
        g.rqmc = e.kind == PE_ENGINE_RQMC ? 1u : 0u;
        g.antithetic = e.antithetic ? 1u : 0u;
        g.domain = domain;
        return g;
    };
    t.train = engine(te, 1u);   // SamplingDomain::LsmTrain
    t.value = engine(ve, 0u);   // SamplingDomain::Valuation
    // EarlyExerciseDiagnostics inputs: an ex-dividend time equal to an
    // exercise time (compile/simulation.rs:437-441), the training engine's
    // RQMC checksums over the exercise-date dimensions
    p.lsm_collisions.assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
        const double et = static_cast<double>(p.obs_days[i]) / 365.0;
        for (const auto& e : p.dividends.timeline)
            if (e.ex_time == et) p.lsm_collisions[i] = 1;
    }
    p.train_diag = pe_mc_diagnostics{};
    if (te.kind == PE_ENGINE_RQMC) {
        p.train_diag.has_checksums = 1;
        rqmc_checksums(te.seed, te.scrambles, static_cast<uint32_t>(n), p.train_diag.direction_checksum,
                       p.train_diag.scramble_checksum);
    }
    t.train_paths = train_paths;
    t.n_dates = static_cast<uint32_t>(n);
    t.call = p.req.product.side == PE_CALL ? 1u : 0u;
    t.zero_variance = p.total_variance == 0.0 ? 1u : 0u;
    t.single_terminal = n == 1 ? 1u : 0u;
    t.basis = l.method == PE_LSM_MLP ? kLsmMlp : l.method == PE_LSM_POLY_LAGUERRE ? kLsmLaguerre : kLsmMonomial;
    t.columns = l.max_degree + 1;
    t.strike = p.req.product.strike;
    t.notional = p.req.product.notional;
    t.s0 = p.spot;
    t.vol = p.vol;
    t.abs_tol = l.abs_rank_tolerance;
    t.rel_tol = l.rel_rank_tolerance;
    t.itm_tol = l.itm_abs_tolerance;
    t.inv_strike = 1.0 / t.strike;
    t.f_vol = static_cast<float>(p.vol);
    t.f_strike = static_cast<float>(t.strike);
    t.f_s0 = static_cast<float>(p.spot);
    t.f_inv_strike = static_cast<float>(t.inv_strike);
    t.risk = p.risk ? 1u : 0u;
    t.request_gamma = p.request_gamma ? 1u : 0u;
    t.gamma_bump = p.gamma_bump;
    t.spot_bump = p.spot_bump;
    t.vol_bump = p.vol_bump;
}

void compile_host_plan(HostPlan& p, const pe_request& r, const pe_exec_policy* policy) {
    static std::atomic<uint64_t> next_id{1};
    p.id = next_id.fetch_add(1);
    p.req = r;
    const pe_product& src = r.product;
    if (src.date_count) {
        if (!src.dates) invalid("product dates missing");
        p.dates.assign(src.dates, src.dates + src.date_count);
        if (src.weights) p.weights.assign(src.weights, src.weights + src.date_count);
        if (src.fixings) p.fixings.assign(src.fixings, src.fixings + src.date_count);
    }
    p.req.product.dates = p.dates.empty() ? nullptr : p.dates.data();
    p.req.product.weights = p.weights.empty() ? nullptr : p.weights.data();
    p.req.product.fixings = p.fixings.empty() ? nullptr : p.fixings.data();
    p.graph = src.kind == PE_PRODUCT_GRAPH;
    if (p.graph) {
        if (src.graph_node_count && src.graph_nodes)
            p.graph_nodes.assign(src.graph_nodes, src.graph_nodes + src.graph_node_count);
        if (src.graph_cashflow_count && src.graph_cashflows)
            p.graph_cashflows.assign(src.graph_cashflows, src.graph_cashflows + src.graph_cashflow_count);
        if (src.graph_underlying_count && src.graph_underlyings)
            p.graph_underlyings.assign(src.graph_underlyings, src.graph_underlyings + src.graph_underlying_count);
    }
    p.req.product.graph_nodes = p.graph_nodes.empty() ? nullptr : p.graph_nodes.data();
    p.req.product.graph_cashflows = p.graph_cashflows.empty() ? nullptr : p.graph_cashflows.data();
    p.req.product.graph_underlyings = p.graph_underlyings.empty() ? nullptr : p.graph_underlyings.data();
    const pe_product& pr = p.req.product;
    if (policy && policy->reduction_block_size) p.block_size = policy->reduction_block_size;

    const pe_market& m = r.market;
    if (!std::isfinite(m.spot) || m.spot <= 0.0) invalid("spot must be positive");
    p.spot = m.spot;
    p.discount_curve.init(m.discount_curve);
    p.dividend_curve.init(m.dividend_curve);
    if (m.dividend_count) {
        if (r.model.kind == PE_MODEL_LOCAL_VOL) unsupported("discrete dividends under local volatility");
        if (!m.dividends) invalid("dividend array missing");
        p.dividends.compile(p.spot, m.dividends, m.dividend_count, &p.discount_curve, &p.dividend_curve);
    }

    p.continuous = pr.kind == PE_PRODUCT_BARRIER && pr.continuous;
    if (p.continuous && r.model.kind == PE_MODEL_LOCAL_VOL) unsupported("continuous barriers under local volatility");
    if (p.graph) {
        // the tape fixes the observation dates, the expiry (last observation)
        // and the discounting date (last payment)
        const bool greeks = r.risk.delta || r.risk.vega || r.risk.gamma;
        if (greeks && r.model.kind == PE_MODEL_LOCAL_VOL) unsupported("pathwise Greeks of graph products under local volatility");
        compile_graph(pr, [&](int32_t day) { return std::exp(p.discount_curve.log_discount(static_cast<double>(day) / 365.0)); },
                      p.graph_tape);
        // Greeks keep each register's two tangents beside its value in LDS
        if (greeks && p.graph_tape.regs > kGraphMaxRiskRegs)
            unsupported("pathwise Greeks of a graph product needing more than " + std::to_string(kGraphMaxRiskRegs) +
                        " registers");
        // local-vol paths carry no discrete dividends: no pre-dividend spot
        if (r.model.kind == PE_MODEL_LOCAL_VOL && !p.graph_tape.pre_days.empty())
            unsupported("pre-dividend graph observation under local volatility");
        p.req.product.expiry_day = p.graph_tape.expiry_day;
        p.req.product.payment_day = p.graph_tape.last_payment;
    } else if (!(pr.strike > 0.0) || !(pr.notional > 0.0)) {
        invalid("strike and notional must be positive");
    }
    for (size_t i = 1; i < p.dates.size(); ++i)
        if (!(p.dates[i] > p.dates[i - 1])) invalid("product dates must increase strictly");
    const std::vector<int32_t> contractual = p.graph ? p.graph_tape.days : fixing_days(pr);
    if (contractual.empty()) unsupported("fully fixed payoffs");
    // Observation points (compile/simulation.rs:210-265): the contractual
    // dates and, for a continuous barrier, the dividend ex-times up to the
    // end of monitoring; equal times keep the contractual date.
    std::vector<std::pair<double, int32_t>> points;
    for (int32_t d : contractual) points.push_back({static_cast<double>(d) / 365.0, d});
    double monitoring_end = 0.0;
    if (p.continuous) {
        if (p.dates.empty()) invalid("barrier needs monitoring dates");
        monitoring_end = static_cast<double>(p.dates.back()) / 365.0;
        std::vector<std::pair<double, int32_t>> kept;
        for (const auto& pt : points)
            if (pt.first > 0.0 || pt.second == pr.expiry_day) kept.push_back(pt);
        points.swap(kept);
        for (const auto& e : p.dividends.timeline)
            if (e.ex_time > 0.0 && e.ex_time <= monitoring_end) points.push_back({e.ex_time, kNoDay});
        std::stable_sort(points.begin(), points.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        std::vector<std::pair<double, int32_t>> unique;
        for (const auto& pt : points)
            if (unique.empty() || unique.back().first != pt.first) unique.push_back(pt);
        points.swap(unique);
    }
    p.obs_days.clear();
    p.obs_times.clear();
    for (const auto& pt : points) {
        p.obs_times.push_back(pt.first);
        p.obs_days.push_back(pt.second);
    }
    p.expiry = static_cast<double>(pr.expiry_day) / 365.0;
    p.forward = p.forward_at(p.expiry);
    p.discount = std::exp(p.discount_curve.log_discount(static_cast<double>(pr.payment_day) / 365.0));

    auto obs_index = [&](int32_t day) -> int32_t {
        for (size_t i = 0; i < p.obs_days.size(); ++i)
            if (p.obs_days[i] == day) return static_cast<int32_t>(i);
        invalid("missing observation");
    };
    p.expiry_obs = obs_index(pr.expiry_day);

    // Product descriptor and per-observation fixing weights.
    ProductDesc& d = p.product;
    std::memset(&d, 0, sizeof(d));
    d.call = pr.side == PE_CALL ? 1 : 0;
    d.strike = pr.strike;
    d.notional = pr.notional;
    d.discount = p.discount;
    d.smoothing = r.risk.smoothing ? 1 : 0;
    d.half_width = r.risk.smoothing_half_width;
    if (d.smoothing && !(d.half_width > 0.0)) invalid("smoothing half width must be positive");
    d.f_strike = static_cast<float>(d.strike);
    d.f_half_width = static_cast<float>(d.half_width);
    d.terminal_obs = -1;
    p.obs_weight.assign(p.obs_days.size(), 0.0);
    switch (pr.kind) {
        case PE_PRODUCT_EUROPEAN:
            d.kind = kEuropean;
            d.terminal_obs = obs_index(pr.expiry_day);
            break;
        case PE_PRODUCT_DIGITAL:
            d.kind = kDigital;
            d.payout = pr.payout;
            d.payout_asset = pr.payout_is_asset ? 1 : 0;
            d.terminal_obs = obs_index(pr.expiry_day);
            break;
        case PE_PRODUCT_ASIAN: {
            d.kind = kAsian;
            if (!pr.weights) invalid("Asian weights missing");
            // The tape adds fixings in date order from 0.0; known fixings must
            // precede future ones so their sum is a prefix.
            double prefix = 0.0;
            bool seen_unknown = false;
            for (uint32_t i = 0; i < pr.date_count; ++i) {
                const bool known = pr.fixings && !std::isnan(pr.fixings[i]);
                if (known) {
                    if (seen_unknown) unsupported("known Asian fixing after a future fixing");
                    prefix = prefix + pr.fixings[i] * pr.weights[i];
                } else {
                    if (pr.dates[i] < 0) invalid("past Asian observation without a fixing");
                    seen_unknown = true;
                    p.obs_weight[obs_index(pr.dates[i])] = pr.weights[i];
                }
            }
            d.asian_prefix = prefix;
            break;
        }
        case PE_PRODUCT_BARRIER:
            d.kind = kBarrier;
            d.barrier = pr.barrier;
            d.f_barrier = static_cast<float>(pr.barrier);
            d.down = pr.direction_down ? 1 : 0;
            d.knock_in = pr.knock_in ? 1 : 0;
            d.has_rebate = pr.has_rebate ? 1 : 0;
            d.rebate = pr.has_rebate ? pr.rebate : 0.0;
            d.initial_hit = pr.historical_hit == 1 ? 1 : 0;
            d.continuous = p.continuous ? 1 : 0;
            if (!(pr.barrier > 0.0) || !std::isfinite(pr.barrier)) invalid("barrier must be positive");
            d.terminal_obs = obs_index(pr.expiry_day);
            if (!p.continuous)
                for (uint32_t i = 0; i < pr.date_count; ++i)
                    if (pr.dates[i] >= 0) p.obs_weight[obs_index(pr.dates[i])] = 1.0;
            break;
        case PE_PRODUCT_LOOKBACK:
            d.kind = kLookback;
            d.has_hist_ext = pr.has_historical_extremum ? 1 : 0;
            d.hist_ext = pr.historical_extremum;
            for (uint32_t i = 0; i < pr.date_count; ++i)
                if (pr.dates[i] >= 0) p.obs_weight[obs_index(pr.dates[i])] = 1.0;
            break;
        case PE_PRODUCT_GRAPH:
            d.kind = kGraph;
            break;
        case PE_PRODUCT_AMERICAN:
            // exercise dates are the observations (execute_early_exercise needs
            // every date in the future: discount factors per date)
            p.american = true;
            d.kind = kEuropean;
            for (uint32_t i = 0; i < pr.date_count; ++i)
                if (pr.dates[i] < 0) invalid("exercise dates must be future and distinct");
            break;
        default:
            unsupported("product kind");
    }

    const pe_model& model = r.model;
    if (model.kind == PE_MODEL_BLACK_SCHOLES || model.kind == PE_MODEL_BLACK76) {
        p.vol = model.volatility;
        if (!std::isfinite(p.vol) || p.vol < 0.0) invalid("volatility must be finite and non-negative");
        p.total_variance = p.vol * p.vol * p.expiry;
        for (double t : p.obs_times) p.obs_forwards.push_back(p.forward_at(t));
    } else if (model.kind == PE_MODEL_LOCAL_VOL) {
        p.lv = true;
        const pe_local_variance_grid& g = model.lv;
        if (!g.time_nodes || !g.x_nodes || !g.values) invalid("local-variance grid arrays missing");
        p.lv_t.assign(g.time_nodes, g.time_nodes + g.time_count);
        p.lv_x.assign(g.x_nodes, g.x_nodes + g.x_count);
        p.lv_v.assign(g.values, g.values + static_cast<size_t>(g.time_count) * g.x_count);
        check_sorted(p.lv_t, true, "local-variance time nodes");
        check_sorted(p.lv_x, false, "local-variance log-moneyness nodes");
        if (!std::isfinite(g.floor) || g.floor <= 0.0 || !std::isfinite(g.cap) || g.cap < g.floor)
            invalid("local-variance floor/cap");
        for (double v : p.lv_v)
            if (!std::isfinite(v) || v < g.floor || v > g.cap) invalid("local-variance value outside [floor, cap]");
        if (p.lv_t.front() != 0.0 || p.lv_t.back() != p.expiry)
            invalid("local-variance grid must span [0, expiry]");
        p.lv_cap = g.cap;
        double max_step = 0.0;
        for (size_t i = 0; i + 1 < p.lv_t.size(); ++i) max_step = std::max(max_step, p.lv_t[i + 1] - p.lv_t[i]);
        // LocalVolTimeGrid::compile: union of grid nodes and events, ceil(dt / max) substeps.
        std::vector<double> events = p.lv_t;
        events.insert(events.end(), p.obs_times.begin(), p.obs_times.end());
        p.nodes = compile_time_grid(events, max_step);
        for (double t : p.nodes) p.node_forwards.push_back(p.forward_at(t));
        p.obs_nodes.clear();
        for (double t : p.obs_times) {
            const auto it = std::find(p.nodes.begin(), p.nodes.end(), t);
            if (it == p.nodes.end()) invalid("observation missing from the local-vol grid");
            p.obs_nodes.push_back(static_cast<uint32_t>(it - p.nodes.begin()));
        }
        p.total_variance = p.nodes.back() * p.lv_cap;
        const size_t n = p.nodes.size() - 1;
        std::vector<double> inv_sqrt_dt;
        if (r.engine.brownian_bridge) {
            compile_bridge_program(p.nodes, p.program, inv_sqrt_dt, p.bridge_depth);
            if (p.bridge_depth > static_cast<uint32_t>(kMaxBridgeDepth))
                throw PlanError(PE_ERR_SIZE_LIMIT, "Brownian-bridge tree deeper than the device stack");
        } else {
            inv_sqrt_dt.assign(n, 0.0);
        }
        p.steps.resize(n);
        for (size_t s = 0; s < n; ++s) {
            LvStep& st = p.steps[s];
            st.t = p.nodes[s];
            st.dt = p.nodes[s + 1] - st.t;
            st.sqrt_dt = std::sqrt(st.dt);
            st.ratio = p.node_forwards[s + 1] / p.node_forwards[s];
            st.inv_sqrt_dt = inv_sqrt_dt[s];
            st.forward = p.node_forwards[s];
            st.next_forward = p.node_forwards[s + 1];
            const size_t tc = lower_cell(p.lv_t, st.t);
            st.t_cell = static_cast<int32_t>(tc);
            st.t_weight = (st.t - p.lv_t[tc]) / (p.lv_t[tc + 1] - p.lv_t[tc]);
            st.node_mask = 0;
            st.m_half_dt = static_cast<float>(-0.5 * st.dt);
            st.f_sqrt_dt = static_cast<float>(st.sqrt_dt);
            st.f_t_weight = static_cast<float>(st.t_weight);
            st.f_inv_sqrt_dt = static_cast<float>(st.inv_sqrt_dt);
            st.f_next_forward = static_cast<float>(st.next_forward);
            st.pad32[0] = st.pad32[1] = st.pad32[2] = 0.0f;
        }
        // node-indexed terminal for the LV kernel
        if (d.terminal_obs >= 0) d.terminal_obs = static_cast<int32_t>(p.obs_nodes[d.terminal_obs]);
    } else {
        unsupported("model kind");
    }

    const uint64_t independent = r.engine.kind == PE_ENGINE_PSEUDO ? r.engine.units : r.engine.scrambles;
    if (r.engine.kind != PE_ENGINE_PSEUDO && r.engine.kind != PE_ENGINE_RQMC) invalid("engine kind");
    if (r.engine.units == 0) invalid("zero sampling units");
    if (r.engine.kind == PE_ENGINE_RQMC && (r.engine.scrambles == 0 || r.engine.units > 0xffffffffull))
        invalid("RQMC needs scrambles and at most 2^32 points");
    if (p.total_variance > 0.0 && independent < 2)
        throw PlanError(PE_ERR_INSUFFICIENT_UNITS, "at least two independent sampling units are required");

    p.request_gamma = r.risk.gamma != 0;
    p.request_delta = r.risk.delta != 0;
    p.request_vega = r.risk.vega != 0;
    p.risk = r.risk.delta || r.risk.vega || p.request_gamma;
    if (p.risk && !p.lv && pr.kind == PE_PRODUCT_DIGITAL && !d.smoothing)
        invalid("pathwise risk of a discontinuous payoff needs compact-C2 smoothing");
    if (p.risk && !p.lv && pr.kind == PE_PRODUCT_BARRIER && !p.continuous && !d.smoothing)
        invalid("pathwise risk of a discontinuous payoff needs compact-C2 smoothing");
    const double residual_spot = p.spot - p.dividends.reserve;
    if (p.request_gamma) p.gamma_bump = r.risk.gamma_relative ? p.spot * r.risk.gamma_bump : r.risk.gamma_bump;
    if (p.request_gamma && (!std::isfinite(p.gamma_bump) || !(p.gamma_bump > 0.0) || p.gamma_bump >= residual_spot))
        invalid("Gamma bump must be positive and below the escrowed spot");
    p.spot_bump = p.request_gamma ? p.gamma_bump : std::min(p.spot * kRelativeSpotBump, 0.5 * residual_spot);
    p.vol_bump = p.vol == 0.0 ? 0.0 : std::min(kVolatilityBump, p.vol * 0.5);
    if (p.lv && (p.request_delta || p.request_gamma)) {
        // local_vol_bump_runtimes (crates/pricing/src/engine/risk/pathwise_local_vol.rs:24-69):
        // the runtime recompiled with the market forward at spot -/+ the bump.
        const double down = p.spot - p.spot_bump, up = p.spot + p.spot_bump;
        if (!(down > 0.0)) invalid("the spot bump leaves a non-positive spot");
        std::vector<double> fdn, fup;
        for (double t : p.nodes) {
            fdn.push_back(p.forward_at_spot(t, down));
            fup.push_back(p.forward_at_spot(t, up));
        }
        p.lv_bumps.resize(p.steps.size());
        for (size_t s = 0; s < p.steps.size(); ++s)
            p.lv_bumps[s] = LvBumpStep{fdn[s], fdn[s + 1] / fdn[s], fup[s], fup[s + 1] / fup[s]};
    }
    p.sobol_dims = p.lv ? static_cast<uint32_t>(p.nodes.size() - 1) : static_cast<uint32_t>(p.obs_times.size());
    if (!p.lv) compile_bs_trade(p, monitoring_end);
    if (p.american) compile_lsm_trade(p);
    compile_diagnostics(p, policy, p.sobol_dims);
}

// MonteCarloDiagnostics of the plan (valuation.rs:94-118 fills the same
// fields): the engine identity, RQMC checksums, the execution, AAD-tile and
// checkpoint policies (AadTilePolicy::resolve / CheckpointPolicy::resolve,
// crates/pricing/src/engine/aad/workspace.rs:66-140) and the curve regions
// of the payment and expiry times (compile/simulation.rs:267-276).
void compile_diagnostics(HostPlan& p, const pe_exec_policy* policy, uint32_t dims) {
    const pe_request& r = p.req;
    pe_mc_diagnostics& d = p.diag;
    d = pe_mc_diagnostics{};
    const pe_engine& e = r.engine;
    d.master_seed = e.seed;
    d.estimator = e.kind == PE_ENGINE_RQMC ? PE_ESTIMATOR_RQMC : PE_ESTIMATOR_PSEUDO_MONTE_CARLO;
    if (e.kind == PE_ENGINE_RQMC) {
        d.scramble_count = e.scrambles;
        d.has_checksums = 1;
        rqmc_checksums(e.seed, e.scrambles, dims, d.direction_checksum, d.scramble_checksum);
    }
    d.policy_version = 1;
    d.worker_threads = policy && policy->worker_threads ? policy->worker_threads : 1u;
    d.reduction_block_size = p.block_size;
    const uint64_t tile = r.risk.aad_tile_capacity ? r.risk.aad_tile_capacity : std::min<uint64_t>(256, p.block_size);
    if (tile > p.block_size) invalid("AAD tile capacity exceeds the reduction block size");
    d.aad_tile_policy_version = 1;
    d.aad_tile_capacity = static_cast<uint32_t>(tile);
    d.checkpoint_policy_version = 1;
    d.checkpoint_interval = r.risk.checkpoint_interval ? r.risk.checkpoint_interval : 64u;
    d.antithetic = e.antithetic ? 1 : 0;
    const double expiry_time = p.obs_times.empty() ? 0.0 : static_cast<double>(r.product.expiry_day) / 365.0;
    d.discount_region = p.discount_curve.region(static_cast<double>(r.product.payment_day) / 365.0);
    d.dividend_region = p.dividend_curve.region(expiry_time);
    d.valuation_kind = r.risk.smoothing ? PE_VALUATION_SMOOTHED_SURROGATE : PE_VALUATION_EXACT_CONTRACTUAL;
}

}  // namespace pe
