// ORACLE -- test infrastructure only.  Graph products restated from the
// reference, compiled and evaluated the way its CPU path does:
//   validation          crates/pricing/src/product/graph_product.rs:53-137
//   cash-flow output    crates/pricing/src/engine/compile/simulation.rs:668-723
//   compile             crates/pricing/src/engine/payoff/compile.rs:490-682
//   constant folding    compile.rs:684-778
//   lazy evaluation     crates/pricing/src/engine/payoff/program.rs:10-258 and
//                       crates/pricing/src/engine/payoff/tape.rs:352-470
// The evaluation is the lazy one: a Select evaluates its condition and then
// only the chosen branch, shared nodes are evaluated once per path, and every
// evaluated node must be finite (check_value, tape.rs:856-865).  Written as an
// explicit-stack recursion rather than the reference's call/return program;
// both evaluate the same node set.
#pragma once

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <set>
#include <stdexcept>
#include <vector>

#include "oracle_payoff.hpp"
#include "pe.h"

namespace oracle {

struct GraphPayoff {
    struct GNode {
        pe_graph_opcode op;
        int in[3];
        int arity;
        pe_graph_comparison cmp;
        int obs;          // observation index (spots)
        double value;     // literal / folded value or half width
        bool constant;
    };
    std::vector<GNode> nodes;   // compiled, in topological order
    int output = -1;
    std::vector<int32_t> days, pre_days;
    int32_t last_payment = 0, expiry_day = 0;

    static int arity(pe_graph_opcode op) {
        switch (op) {
            case PE_GRAPH_LITERAL: case PE_GRAPH_TERMINAL_SPOT: case PE_GRAPH_PRE_DIVIDEND_SPOT: return 0;
            case PE_GRAPH_INDICATOR: case PE_GRAPH_NEGATE: case PE_GRAPH_SMOOTH_INDICATOR:
            case PE_GRAPH_LOG: case PE_GRAPH_EXP: return 1;
            case PE_GRAPH_SELECT: return 3;
            default: return 2;
        }
    }
    static bool test(pe_graph_comparison c, double a, double b) {
        switch (c) {
            case PE_GRAPH_LESS: return a < b;
            case PE_GRAPH_LESS_EQUAL: return a <= b;
            case PE_GRAPH_EQUAL: return a == b;
            case PE_GRAPH_NOT_EQUAL: return a != b;
            case PE_GRAPH_GREATER_EQUAL: return a >= b;
            default: return a > b;
        }
    }
    // execute_opcode for a non-Select operation on operand values x[]
    static double apply(const GNode& g, const double* x) {
        const Smooth sm{g.value};
        double v = 0.0, d1, d2;
        switch (g.op) {
            case PE_GRAPH_ADD: return x[0] + x[1];
            case PE_GRAPH_SUBTRACT: return x[0] - x[1];
            case PE_GRAPH_MULTIPLY: return x[0] * x[1];
            case PE_GRAPH_DIVIDE: return x[0] / x[1];
            case PE_GRAPH_MINIMUM: return x[0] <= x[1] ? x[0] : x[1];
            case PE_GRAPH_MAXIMUM: return x[0] >= x[1] ? x[0] : x[1];
            case PE_GRAPH_INDICATOR: return x[0] >= 0.0 ? 1.0 : 0.0;
            case PE_GRAPH_NEGATE: return -x[0];
            case PE_GRAPH_BARRIER_HIT_STATE: return x[0] + (1.0 - x[0]) * x[1];
            case PE_GRAPH_SMOOTH_MAXIMUM: sm.maximum(x[0], x[1], &v, &d1, &d2); return v;
            case PE_GRAPH_SMOOTH_MINIMUM: {
                // SmoothMinimum: left - positive_part(left - right) inside the band
                const double d = x[0] - x[1];
                if (d >= sm.h) return x[1];
                if (d <= -sm.h) return x[0];
                double c;
                sm.positive_part(d, &v, &d1, &c);
                return x[0] - v;
            }
            case PE_GRAPH_SMOOTH_INDICATOR: sm.indicator(x[0], &v, &d1); return v;
            case PE_GRAPH_LOG: return std::log(x[0]);
            case PE_GRAPH_EXP: return std::exp(x[0]);
            case PE_GRAPH_COMPARE: return test(g.cmp, x[0], x[1]) ? 1.0 : 0.0;
            default: throw std::logic_error("graph opcode");
        }
    }

    void compile(const pe_product& p, const std::function<double(int32_t)>& discount) {
        if (!p.graph_nodes || !p.graph_node_count) throw std::invalid_argument("graph without nodes");
        if (!p.graph_cashflows || !p.graph_cashflow_count) throw std::invalid_argument("graph without cash flows");
        if (!p.graph_underlyings || !p.graph_underlying_count) throw std::invalid_argument("graph underlyings");
        std::set<uint32_t> declared(p.graph_underlyings, p.graph_underlyings + p.graph_underlying_count);
        if (declared.size() != p.graph_underlying_count) throw std::invalid_argument("duplicate graph underlying");
        struct Src {
            pe_graph_opcode op;
            uint32_t in[3];
            pe_graph_comparison cmp;
            int32_t day;
            double value;
        };
        std::map<uint32_t, Src> src;
        bool branching = false;
        for (uint32_t i = 0; i < p.graph_node_count; ++i) {
            const pe_graph_node& n = p.graph_nodes[i];
            Src s{n.opcode, {n.operands[0], n.operands[1], n.operands[2]}, n.comparison, n.observation_day, n.value};
            if (n.opcode == PE_GRAPH_LITERAL && !std::isfinite(n.value)) throw std::invalid_argument("literal");
            if ((n.opcode == PE_GRAPH_TERMINAL_SPOT || n.opcode == PE_GRAPH_PRE_DIVIDEND_SPOT)) {
                if (!declared.count(n.underlying_id)) throw std::invalid_argument("undeclared underlying");
                if (n.underlying_id != p.graph_underlyings[0]) throw std::invalid_argument("foreign underlying");
            }
            branching = branching || n.opcode == PE_GRAPH_SELECT;
            if (!src.emplace(n.id, s).second) throw std::invalid_argument("duplicate node id");
        }
        for (const auto& kv : src)
            for (int j = 0; j < arity(kv.second.op); ++j)
                if (!src.count(kv.second.in[j])) throw std::invalid_argument("unknown operand");
        // topological order: smallest ready id first
        auto topo = [&]() {
            std::map<uint32_t, int> deg;
            std::map<uint32_t, std::vector<uint32_t>> out;
            for (const auto& kv : src) {
                deg[kv.first] = arity(kv.second.op);
                for (int j = 0; j < arity(kv.second.op); ++j) out[kv.second.in[j]].push_back(kv.first);
            }
            std::set<uint32_t> ready;
            for (const auto& kv : deg)
                if (kv.second == 0) ready.insert(kv.first);
            std::vector<uint32_t> order;
            while (!ready.empty()) {
                const uint32_t id = *ready.begin();
                ready.erase(ready.begin());
                order.push_back(id);
                for (uint32_t d : out[id])
                    if (--deg[d] == 0) ready.insert(d);
            }
            if (order.size() != src.size()) throw std::invalid_argument("cycle");
            return order;
        };
        std::vector<uint32_t> order = topo();
        // latest observation per node; payments must not precede it
        std::map<uint32_t, int32_t> latest;
        for (uint32_t id : order) {
            const Src& s = src[id];
            int32_t d = INT32_MIN;
            if (s.op == PE_GRAPH_TERMINAL_SPOT || s.op == PE_GRAPH_PRE_DIVIDEND_SPOT) d = s.day;
            for (int j = 0; j < arity(s.op); ++j) d = std::max(d, latest[s.in[j]]);
            latest[id] = d;
        }
        last_payment = INT32_MIN;
        for (uint32_t i = 0; i < p.graph_cashflow_count; ++i) {
            const pe_graph_cashflow& c = p.graph_cashflows[i];
            if (!src.count(c.output_node)) throw std::invalid_argument("unknown output");
            if (latest[c.output_node] != INT32_MIN && latest[c.output_node] > c.payment_day)
                throw std::invalid_argument("cash flow depends on an observation after its payment");
            last_payment = std::max(last_payment, c.payment_day);
        }
        uint32_t out_id = p.graph_cashflows[0].output_node;
        if (p.graph_cashflow_count > 1) {
            uint32_t next = src.rbegin()->first + 1;
            const double final_df = discount(last_payment);
            uint32_t sum = 0;
            for (uint32_t i = 0; i < p.graph_cashflow_count; ++i) {
                const pe_graph_cashflow& c = p.graph_cashflows[i];
                uint32_t v = c.output_node;
                if (c.payment_day != last_payment) {
                    const uint32_t lit = next++;
                    src[lit] = Src{PE_GRAPH_LITERAL, {0, 0, 0}, PE_GRAPH_LESS, 0, discount(c.payment_day) / final_df};
                    v = next++;
                    src[v] = Src{PE_GRAPH_MULTIPLY, {c.output_node, lit, 0}, PE_GRAPH_LESS, 0, 0.0};
                }
                if (i == 0) {
                    sum = v;
                } else {
                    const uint32_t a = next++;
                    src[a] = Src{PE_GRAPH_ADD, {sum, v, 0}, PE_GRAPH_LESS, 0, 0.0};
                    sum = a;
                }
            }
            out_id = sum;
            order = topo();
        }
        // folding
        std::map<uint32_t, std::pair<bool, double>> fold;
        for (uint32_t id : order) {
            const Src& s = src[id];
            std::pair<bool, double> f{false, 0.0};
            if (s.op == PE_GRAPH_LITERAL) {
                f = {true, s.value};
            } else if (s.op == PE_GRAPH_SELECT) {
                if (fold[s.in[0]].first) {
                    const auto& pick = fold[fold[s.in[0]].second != 0.0 ? s.in[1] : s.in[2]];
                    if (pick.first) f = pick;
                }
            } else if (s.op != PE_GRAPH_TERMINAL_SPOT && s.op != PE_GRAPH_PRE_DIVIDEND_SPOT) {
                double x[3] = {0, 0, 0};
                bool all = true;
                for (int j = 0; j < arity(s.op); ++j) {
                    all = all && fold[s.in[j]].first;
                    x[j] = fold[s.in[j]].second;
                }
                if (all) {
                    GNode g{s.op, {0, 0, 0}, 0, s.cmp, -1, s.value, false};
                    const double v = apply(g, x);
                    if (std::isfinite(v)) f = {true, v};
                    else if (!branching) throw std::invalid_argument("non-finite constant fold");
                }
            }
            fold[id] = f;
        }
        // output cone
        std::set<uint32_t> reach;
        std::vector<uint32_t> stack{out_id};
        while (!stack.empty()) {
            const uint32_t id = stack.back();
            stack.pop_back();
            if (!reach.insert(id).second) continue;
            for (int j = 0; j < arity(src[id].op); ++j) stack.push_back(src[id].in[j]);
        }
        std::set<int32_t> ds, pds;
        for (uint32_t id : reach) {
            const Src& s = src[id];
            if (s.op == PE_GRAPH_TERMINAL_SPOT || s.op == PE_GRAPH_PRE_DIVIDEND_SPOT) {
                if (s.day < 0) throw std::invalid_argument("graph observation before valuation");
                ds.insert(s.day);
                if (s.op == PE_GRAPH_PRE_DIVIDEND_SPOT) pds.insert(s.day);
            }
        }
        days.assign(ds.begin(), ds.end());
        pre_days.assign(pds.begin(), pds.end());
        expiry_day = days.empty() ? last_payment : days.back();
        std::map<uint32_t, int> slot;
        nodes.clear();
        for (uint32_t id : order) {
            if (!reach.count(id)) continue;
            const Src& s = src[id];
            GNode g{s.op, {0, 0, 0}, arity(s.op), s.cmp, -1, s.value, fold[id].first};
            if (g.constant) g.value = fold[id].second;
            for (int j = 0; j < g.arity; ++j) g.in[j] = slot.at(s.in[j]);
            if (s.op == PE_GRAPH_TERMINAL_SPOT || s.op == PE_GRAPH_PRE_DIVIDEND_SPOT)
                g.obs = static_cast<int>(std::lower_bound(days.begin(), days.end(), s.day) - days.begin());
            slot[id] = static_cast<int>(nodes.size());
            nodes.push_back(g);
        }
        output = slot.at(out_id);
    }

    // Lazy evaluation of the output on one path: post[i] / pre[i] are the
    // spots at observation i; has_pre[i] marks a pre-dividend spot.
    double evaluate(const double* post, const double* pre, const char* has_pre) const {
        std::vector<double> v;
        std::vector<char> done;
        return run(post, pre, has_pre, v, done);
    }

    // evaluate_single_with_observation_adjoints (tape.rs:484-600): the lazy
    // evaluation, then the reverse sweep over the evaluated nodes in reverse
    // order with reverse_opcode's pullbacks (tape.rs:602-780); post_adj[i] /
    // pre_adj[i] (zeroed by the caller) receive d output / d spot at
    // observation i.  A non-finite adjoint is an error (add_adjoint).
    double evaluate_adjoints(const double* post, const double* pre, const char* has_pre, double* post_adj,
                             double* pre_adj) const {
        std::vector<double> v;
        std::vector<char> done;
        const double value = run(post, pre, has_pre, v, done);
        std::vector<double> adj(nodes.size(), 0.0);
        adj[output] = 1.0;
        auto add = [&](int slot, double c) {
            const double u = adj[slot] + c;
            if (!std::isfinite(u)) throw std::runtime_error("non-finite graph runtime adjoint");
            adj[slot] = u;
        };
        for (int id = static_cast<int>(nodes.size()) - 1; id >= 0; --id) {
            if (!done[id]) continue;
            const GNode& g = nodes[id];
            const double o = adj[id];
            if (g.constant) continue;   // a literal or folded value has no inputs
            const double x = g.arity > 0 ? v[g.in[0]] : 0.0, y = g.arity > 1 ? v[g.in[1]] : 0.0;
            const Smooth sm{g.value};
            switch (g.op) {
                case PE_GRAPH_TERMINAL_SPOT: post_adj[g.obs] += o; break;
                case PE_GRAPH_PRE_DIVIDEND_SPOT: pre_adj[g.obs] += o; break;
                case PE_GRAPH_SELECT: add(x != 0.0 ? g.in[1] : g.in[2], o); break;
                case PE_GRAPH_ADD: add(g.in[0], o); add(g.in[1], o); break;
                case PE_GRAPH_SUBTRACT: add(g.in[0], o); add(g.in[1], -o); break;
                case PE_GRAPH_MULTIPLY: add(g.in[0], o * y); add(g.in[1], o * x); break;
                case PE_GRAPH_DIVIDE: add(g.in[0], o / y); add(g.in[1], denominator_adjoint(o, x, y)); break;
                case PE_GRAPH_MINIMUM: add(x <= y ? g.in[0] : g.in[1], o); break;
                case PE_GRAPH_MAXIMUM: add(x >= y ? g.in[0] : g.in[1], o); break;
                case PE_GRAPH_NEGATE: add(g.in[0], -o); break;
                case PE_GRAPH_LOG: add(g.in[0], o / x); break;
                case PE_GRAPH_EXP: add(g.in[0], o * v[id]); break;
                case PE_GRAPH_BARRIER_HIT_STATE: add(g.in[0], o * (1.0 - y)); add(g.in[1], o * (1.0 - x)); break;
                case PE_GRAPH_SMOOTH_MAXIMUM: {
                    double val, lf, rf;
                    sm.maximum(x, y, &val, &lf, &rf);
                    add(g.in[0], o * lf);
                    add(g.in[1], o * rf);
                    break;
                }
                case PE_GRAPH_SMOOTH_MINIMUM: {
                    // SmoothingPolicy::minimum (product/smoothing.rs:123-145)
                    const double d = x - y;
                    double lf, rf;
                    if (d >= sm.h) { lf = 0.0; rf = 1.0; }
                    else if (d <= -sm.h) { lf = 1.0; rf = 0.0; }
                    else {
                        double pv, pf, comp;
                        sm.positive_part(d, &pv, &pf, &comp);
                        lf = comp;
                        rf = pf;
                    }
                    add(g.in[0], o * lf);
                    add(g.in[1], o * rf);
                    break;
                }
                case PE_GRAPH_SMOOTH_INDICATOR: {
                    double val, first;
                    sm.indicator(x, &val, &first);
                    add(g.in[0], o * first);
                    break;
                }
                default: break;   // Compare, Indicator, Literal: no derivative
            }
        }
        return value;
    }

    // division_denominator_adjoint (tape.rs:783-805): -seed n / d^2 without
    // losing a finite value when the product or the square leaves the normal
    // range (fractions in [1, 2) and one final scaling).
    static double denominator_adjoint(double seed, double n, double d) {
        const double product = -seed * n, square = d * d;
        if (std::isnormal(square) && std::isnormal(product)) return product / square;
        if (seed == 0.0 || n == 0.0) return product;
        auto parts = [](double x, int* e) {
            const double m = std::frexp(x, e);   // m in [0.5, 1)
            *e -= 1;
            return 2.0 * m;                      // in [1, 2)
        };
        int es, en, ed;
        const double fs = parts(seed, &es), fn = parts(n, &en), fd = parts(d, &ed);
        return std::ldexp(-fs * fn / (fd * fd), es + en - 2 * ed);
    }

    double run(const double* post, const double* pre, const char* has_pre, std::vector<double>& v,
               std::vector<char>& done) const {
        v.assign(nodes.size(), 0.0);
        done.assign(nodes.size(), 0);
        struct Frame { int id, stage; };
        std::vector<Frame> st{{output, 0}};
        while (!st.empty()) {
            const int id = st.back().id;
            const int stage = st.back().stage;
            const GNode& g = nodes[id];
            if (done[id]) { st.pop_back(); continue; }
            double x = 0.0;
            if (g.constant) {
                x = g.value;
            } else if (g.op == PE_GRAPH_TERMINAL_SPOT) {
                x = post[g.obs];
            } else if (g.op == PE_GRAPH_PRE_DIVIDEND_SPOT) {
                if (!has_pre[g.obs]) throw std::runtime_error("missing pre-dividend observation");
                x = pre[g.obs];
            } else if (g.op == PE_GRAPH_SELECT) {
                if (stage == 0) {  // the condition first
                    st.back().stage = 1;
                    if (!done[g.in[0]]) st.push_back({g.in[0], 0});
                    continue;
                }
                const int pick = v[g.in[0]] != 0.0 ? g.in[1] : g.in[2];
                if (stage == 1) {  // then only the chosen branch
                    st.back().stage = 2;
                    if (!done[pick]) st.push_back({pick, 0});
                    continue;
                }
                x = v[pick];
            } else {
                if (stage < g.arity) {
                    st.back().stage = stage + 1;
                    if (!done[g.in[stage]]) st.push_back({g.in[stage], 0});
                    continue;
                }
                const double in[3] = {v[g.in[0]], g.arity > 1 ? v[g.in[1]] : 0.0, 0.0};
                x = apply(g, in);
            }
            if (!std::isfinite(x)) throw std::runtime_error("non-finite graph runtime value");
            v[id] = x;
            done[id] = 1;
            st.pop_back();
        }
        return v[output];
    }
};

}  // namespace oracle
