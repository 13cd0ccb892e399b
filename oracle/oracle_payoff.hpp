// ORACLE -- test infrastructure only.  A small SSA payoff tape restating the
// reference's product graphs and tape semantics:
//   product graphs     crates/pricing/src/engine/payoff/compile.rs:35-474
//   forward opcodes    crates/pricing/src/engine/payoff/tape.rs:953-1088
//   reverse opcodes    crates/pricing/src/engine/payoff/tape.rs:594-781
//   compact C2         crates/pricing/src/product/smoothing.rs:5-198
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "pe.h"

namespace oracle {

struct Smooth {  // CompactC2Smoothing
    double h;

    static double quintic(double u) { return u * u * u * (10.0 + u * (-15.0 + 6.0 * u)); }
    double transition(double input) const {
        const double full = 2.0 * h;
        if (std::isfinite(full)) return (input + h) / full;
        if (input < 0.0) return 0.5 * ((input + h) / h);
        return 0.5 * (input / h + 1.0);
    }
    // interior_indicator: returns value, first, second; complement in *comp.
    void interior(double input, double u, double* value, double* first, double* second, double* comp) const {
        const double hit = quintic(u);
        const double safety = 1.0 - hit;
        double z, v, c, sign;
        if (safety <= 32.0 * DBL_EPSILON) {
            z = transition(-input);
            const double tail = quintic(z);
            v = 1.0 - tail; c = tail; sign = -1.0;
        } else {
            z = u; v = hit; c = safety; sign = 1.0;
        }
        const double z2 = z * z, z3 = z2 * z, omz = 1.0 - z;
        *first = 15.0 * z2 * omz * omz / h;
        const double numerator = sign * (120.0 * z3 - 180.0 * z2 + 60.0 * z);
        const double denominator = 4.0 * h * h;
        *second = std::isnormal(denominator) ? numerator / denominator : ((0.25 * numerator) / h) / h;
        *value = v;
        *comp = c;
    }
    void indicator(double input, double* value, double* first) const {
        double comp;
        indicator_with_complement(input, value, first, &comp);
    }
    // indicator_with_complement (smoothing.rs:34-50): the hit weight, its first
    // derivative and the safety weight 1 - hit computed without cancellation.
    void indicator_with_complement(double input, double* value, double* first, double* comp) const {
        if (input <= -h) { *value = 0.0; *first = 0.0; *comp = 1.0; return; }
        if (input >= h) { *value = 1.0; *first = 0.0; *comp = 0.0; return; }
        double second;
        interior(input, transition(input), value, first, &second, comp);
    }
    void positive_part(double input, double* value, double* first, double* comp) const {
        if (input <= -h) { *value = 0.0; *first = 0.0; *comp = 1.0; return; }
        if (input >= h) { *value = input; *first = 1.0; *comp = 0.0; return; }
        const double u = transition(input);
        const double u2 = u * u, u4 = u2 * u2;
        double iv, ifirst, isecond, icomp;
        interior(input, u, &iv, &ifirst, &isecond, &icomp);
        const double scaled = 2.0 * h * u4;
        const double poly = 2.5 + u * (-3.0 + u);
        *value = std::isnormal(scaled) ? scaled * poly : (2.0 * u4 * poly) * h;
        *first = iv;
        *comp = icomp;
    }
    void maximum(double l, double r, double* value, double* lf, double* rf) const {
        const double d = l - r;
        if (d >= h) { *value = l; *lf = 1.0; *rf = 0.0; return; }
        if (d <= -h) { *value = r; *lf = 0.0; *rf = 1.0; return; }
        double pv, pf, comp;
        positive_part(d, &pv, &pf, &comp);
        *value = r + pv; *lf = pf; *rf = comp;
    }
};

enum class Op { Literal, Spot, PreSpot, Add, Sub, Mul, Max, Min, Ind, SmoothInd, SmoothMax, HitState };

struct Node {
    Op op;
    int a = -1, b = -1;
    double literal = 0.0;
    int obs = -1;  // observation index for Spot / PreSpot
};

struct Tape {
    std::vector<Node> nodes;
    int output = -1;
    Smooth smooth{1.0};

    int push(Node n) { nodes.push_back(n); return static_cast<int>(nodes.size()) - 1; }
    int lit(double v) { Node n{Op::Literal}; n.literal = v; return push(n); }
    int spot(int obs) { Node n{Op::Spot}; n.obs = obs; return push(n); }
    int pre_spot(int obs) { Node n{Op::PreSpot}; n.obs = obs; return push(n); }
    int bin(Op op, int a, int b) { Node n{op}; n.a = a; n.b = b; return push(n); }
    int un(Op op, int a) { Node n{op}; n.a = a; return push(n); }

    // Forward (execute_opcode) with observation post spots and pre-dividend spots.
    void forward(const double* post, const double* pre, std::vector<double>& v) const {
        v.assign(nodes.size(), 0.0);
        for (size_t i = 0; i < nodes.size(); ++i) {
            const Node& n = nodes[i];
            double x = 0.0;
            switch (n.op) {
                case Op::Literal: x = n.literal; break;
                case Op::Spot: x = post[n.obs]; break;
                case Op::PreSpot: x = pre[n.obs]; break;
                case Op::Add: x = v[n.a] + v[n.b]; break;
                case Op::Sub: x = v[n.a] - v[n.b]; break;
                case Op::Mul: x = v[n.a] * v[n.b]; break;
                case Op::Max: x = v[n.a] >= v[n.b] ? v[n.a] : v[n.b]; break;
                case Op::Min: x = v[n.a] <= v[n.b] ? v[n.a] : v[n.b]; break;
                case Op::Ind: x = v[n.a] >= 0.0 ? 1.0 : 0.0; break;
                case Op::SmoothInd: { double f; smooth.indicator(v[n.a], &x, &f); break; }
                case Op::SmoothMax: { double lf, rf; smooth.maximum(v[n.a], v[n.b], &x, &lf, &rf); break; }
                case Op::HitState: x = v[n.a] + (1.0 - v[n.a]) * v[n.b]; break;
            }
            if (!std::isfinite(x)) throw std::runtime_error("non-finite payoff value");
            v[i] = x;
        }
    }

    // Reverse (reverse_opcode) from d(output) = 1; per-observation adjoints of
    // post spots and pre-dividend spots are summed into post_adj / pre_adj.
    void reverse(const std::vector<double>& v, double* post_adj, double* pre_adj) const {
        std::vector<double> adj(nodes.size(), 0.0);
        adj[output] = 1.0;
        for (size_t k = nodes.size(); k-- > 0;) {
            const Node& n = nodes[k];
            const double o = adj[k];
            switch (n.op) {
                case Op::Literal: break;
                case Op::Spot: post_adj[n.obs] += o; break;
                case Op::PreSpot: pre_adj[n.obs] += o; break;
                case Op::Add: adj[n.a] += o; adj[n.b] += o; break;
                case Op::Sub: adj[n.a] += o; adj[n.b] += -o; break;
                case Op::Mul: adj[n.a] += o * v[n.b]; adj[n.b] += o * v[n.a]; break;
                case Op::Max: if (v[n.a] >= v[n.b]) adj[n.a] += o; else adj[n.b] += o; break;
                case Op::Min: if (v[n.a] <= v[n.b]) adj[n.a] += o; else adj[n.b] += o; break;
                case Op::Ind: break;
                case Op::SmoothInd: { double val, f; smooth.indicator(v[n.a], &val, &f); adj[n.a] += o * f; break; }
                case Op::SmoothMax: {
                    double val, lf, rf;
                    smooth.maximum(v[n.a], v[n.b], &val, &lf, &rf);
                    adj[n.a] += o * lf; adj[n.b] += o * rf;
                    break;
                }
                case Op::HitState: {
                    const double p = v[n.a], w = v[n.b];
                    adj[n.a] += o * (1.0 - w);
                    adj[n.b] += o * (1.0 - p);
                    break;
                }
            }
        }
    }
};

// Observation schedule helper: index of the observation at `day`.
struct ObsIndex {
    std::vector<int32_t> days;  // sorted unique
    int of(int32_t day) const {
        for (size_t i = 0; i < days.size(); ++i)
            if (days[i] == day) return static_cast<int>(i);
        throw std::runtime_error("missing observation date");
    }
};

// Builds the product graph.  `jump_obs[i]` marks observations that coincide
// with a dividend ex-date (barrier pre-dividend monitoring, compile.rs:311-339).
inline Tape build_tape(const pe_product& p, const ObsIndex& obs, bool smoothing, double half_width,
                       const std::vector<bool>& jump_obs) {
    Tape t;
    t.smooth.h = half_width;
    const bool call = p.side == PE_CALL;
    switch (p.kind) {
        case PE_PRODUCT_EUROPEAN: {
            const int s = t.spot(obs.of(p.expiry_day));
            const int k = t.lit(p.strike);
            const int sd = call ? t.bin(Op::Sub, s, k) : t.bin(Op::Sub, k, s);
            const int z = t.lit(0.0);
            const int pp = t.bin(Op::Max, sd, z);
            const int n = t.lit(p.notional);
            t.output = t.bin(Op::Mul, pp, n);
            break;
        }
        case PE_PRODUCT_DIGITAL: {
            const int s = t.spot(obs.of(p.expiry_day));
            const int k = t.lit(p.strike);
            const int sd = call ? t.bin(Op::Sub, s, k) : t.bin(Op::Sub, k, s);
            const int ind = smoothing ? t.un(Op::SmoothInd, sd) : t.un(Op::Ind, sd);
            const int pay = t.lit(p.payout);
            const int base = p.payout_is_asset ? t.bin(Op::Mul, s, pay) : pay;
            t.output = t.bin(Op::Mul, ind, base);
            break;
        }
        case PE_PRODUCT_ASIAN: {
            int ws = t.lit(0.0);
            for (uint32_t i = 0; i < p.date_count; ++i) {
                const bool known = p.fixings && !std::isnan(p.fixings[i]);
                const int o = known ? t.lit(p.fixings[i]) : t.spot(obs.of(p.dates[i]));
                const int w = t.lit(p.weights[i]);
                const int wo = t.bin(Op::Mul, o, w);
                ws = t.bin(Op::Add, ws, wo);
            }
            const int k = t.lit(p.strike);
            const int sd = call ? t.bin(Op::Sub, ws, k) : t.bin(Op::Sub, k, ws);
            const int z = t.lit(0.0);
            const int pp = t.bin(Op::Max, sd, z);
            const int n = t.lit(p.notional);
            t.output = t.bin(Op::Mul, pp, n);
            break;
        }
        case PE_PRODUCT_BARRIER: {
            const int k = t.lit(p.strike);
            const int term = t.spot(obs.of(p.expiry_day));
            const int sd = call ? t.bin(Op::Sub, term, k) : t.bin(Op::Sub, k, term);
            const int z = t.lit(0.0);
            const int pp = t.bin(Op::Max, sd, z);
            const int n = t.lit(p.notional);
            const int vanilla = t.bin(Op::Mul, pp, n);
            const int b = t.lit(p.barrier);
            int hit = p.historical_hit == 1 ? t.lit(1.0) : z;
            const bool skip_all = p.continuous && p.historical_hit == 1;
            for (uint32_t i = 0; i < p.date_count && !skip_all; ++i) {
                if (p.dates[i] < 0) continue;
                const int oi = obs.of(p.dates[i]);
                const int s = t.spot(oi);
                int dist = p.direction_down ? t.bin(Op::Sub, b, s) : t.bin(Op::Sub, s, b);
                if (!jump_obs.empty() && jump_obs[oi]) {
                    const int ps = t.pre_spot(oi);
                    const int pd = p.direction_down ? t.bin(Op::Sub, b, ps) : t.bin(Op::Sub, ps, b);
                    dist = smoothing ? t.bin(Op::SmoothMax, pd, dist) : t.bin(Op::Max, pd, dist);
                }
                const int dh = smoothing ? t.un(Op::SmoothInd, dist) : t.un(Op::Ind, dist);
                hit = t.bin(Op::HitState, hit, dh);
            }
            int active;
            if (p.knock_in) {
                active = hit;
            } else {
                const int one = t.lit(1.0);
                active = t.bin(Op::Sub, one, hit);
            }
            const int ap = t.bin(Op::Mul, vanilla, active);
            if (p.has_rebate) {
                const int one = t.lit(1.0);
                const int inactive = t.bin(Op::Sub, one, active);
                const int r = t.lit(p.rebate);
                const int ip = t.bin(Op::Mul, inactive, r);
                t.output = t.bin(Op::Add, ap, ip);
            } else {
                t.output = ap;
            }
            break;
        }
        case PE_PRODUCT_LOOKBACK: {
            std::vector<int32_t> dates;
            for (uint32_t i = 0; i < p.date_count; ++i)
                if (p.dates[i] >= 0) dates.push_back(p.dates[i]);
            size_t start = 0;
            int ext;
            if (p.has_historical_extremum) {
                ext = t.lit(p.historical_extremum);
            } else {
                if (dates.empty()) throw std::invalid_argument("lookback without observations");
                ext = t.spot(obs.of(dates[0]));
                start = 1;
            }
            for (size_t i = start; i < dates.size(); ++i) {
                const int s = t.spot(obs.of(dates[i]));
                ext = call ? t.bin(Op::Max, ext, s) : t.bin(Op::Min, ext, s);
            }
            const int k = t.lit(p.strike);
            const int sd = call ? t.bin(Op::Sub, ext, k) : t.bin(Op::Sub, k, ext);
            const int z = t.lit(0.0);
            const int pp = t.bin(Op::Max, sd, z);
            const int n = t.lit(p.notional);
            t.output = t.bin(Op::Mul, pp, n);
            break;
        }
        default:
            throw std::invalid_argument("unsupported product kind");
    }
    return t;
}

// Sorted unique observation days of a product (CompiledPayoff::terminal_observations,
// tape.rs:229-247), restricted to days on or after valuation.
inline std::vector<int32_t> observation_days(const pe_product& p) {
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
            throw std::invalid_argument("unsupported product kind");
    }
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    return d;
}

}  // namespace oracle
