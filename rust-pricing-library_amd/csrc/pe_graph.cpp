// Graph products (GraphProduct, crates/pricing/src/product/graph_product.rs:40-191)
// compiled for the BS batch kernel's tape interpreter.
//
// The source graph is validated, its cash flows are folded into one output
// (compile_graph_cashflows, crates/pricing/src/engine/compile/simulation.rs:668-723),
// ordered, constant-folded and pruned to the output's cone as the reference's
// compiler does (crates/pricing/src/engine/payoff/compile.rs:490-767).  The
// tape is then scheduled for a path walk: each operation runs right after the
// latest observation it depends on, so path-dependent sums and extrema need a
// few live values instead of one per fixing, and a linear scan assigns the
// live values to at most kGraphMaxRegs registers.
//
// Evaluation is eager (every lane runs the whole tape, Select picks per lane)
// where the reference's program is lazy (program.rs:60-258).  Values agree:
// an untaken branch's value never reaches a taken one.  The reference's
// run-time error -- a non-finite value on any operation it evaluates -- is
// kept exactly by a poison bit per register: the operation's own
// non-finiteness or'ed with its operands' bits, a Select taking only its
// condition's and its chosen branch's.  A set bit on the output is the
// reference's NonFiniteRuntimeValue.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>

#include "pe_plan.hpp"

namespace pe {
namespace {

[[noreturn]] void invalid(const std::string& what) { throw PlanError(PE_ERR_INVALID_ARGUMENT, what); }
[[noreturn]] void unsupported(const std::string& what) { throw PlanError(PE_ERR_UNSUPPORTED, what); }

int operand_count(pe_graph_opcode op) {  // SourceOpcode::operands (graph.rs:170-200)
    switch (op) {
        case PE_GRAPH_LITERAL:
        case PE_GRAPH_TERMINAL_SPOT:
        case PE_GRAPH_PRE_DIVIDEND_SPOT:
            return 0;
        case PE_GRAPH_INDICATOR:
        case PE_GRAPH_NEGATE:
        case PE_GRAPH_SMOOTH_INDICATOR:
        case PE_GRAPH_LOG:
        case PE_GRAPH_EXP:
            return 1;
        case PE_GRAPH_SELECT:
            return 3;
        case PE_GRAPH_ADD:
        case PE_GRAPH_SUBTRACT:
        case PE_GRAPH_MULTIPLY:
        case PE_GRAPH_DIVIDE:
        case PE_GRAPH_MINIMUM:
        case PE_GRAPH_MAXIMUM:
        case PE_GRAPH_BARRIER_HIT_STATE:
        case PE_GRAPH_SMOOTH_MINIMUM:
        case PE_GRAPH_SMOOTH_MAXIMUM:
        case PE_GRAPH_COMPARE:
            return 2;
    }
    invalid("unknown graph opcode");
}

struct Source {
    pe_graph_opcode op;
    uint32_t in[3];
    pe_graph_comparison cmp;
    int32_t day;
    double value;
};

bool compare(pe_graph_comparison c, double a, double b) {  // GraphComparison::test
    switch (c) {
        case PE_GRAPH_LESS: return a < b;
        case PE_GRAPH_LESS_EQUAL: return a <= b;
        case PE_GRAPH_EQUAL: return a == b;
        case PE_GRAPH_NOT_EQUAL: return a != b;
        case PE_GRAPH_GREATER_EQUAL: return a >= b;
        case PE_GRAPH_GREATER: return a > b;
    }
    invalid("unknown graph comparison");
}

// CompactC2Smoothing (crates/pricing/src/product/smoothing.rs:28-198), values only.
struct HostSmooth {
    double h;
    static double quintic(double u) { return u * u * u * (10.0 + u * (-15.0 + 6.0 * u)); }
    double transition(double x) const {
        const double full = 2.0 * h;
        if (std::isfinite(full)) return (x + h) / full;
        if (x < 0.0) return 0.5 * ((x + h) / h);
        return 0.5 * (x / h + 1.0);
    }
    double interior(double x, double u) const {
        const double hit = quintic(u);
        if (1.0 - hit <= 32.0 * 2.220446049250313e-16) return 1.0 - quintic(transition(-x));
        return hit;
    }
    double indicator(double x) const {
        if (x <= -h) return 0.0;
        if (x >= h) return 1.0;
        return interior(x, transition(x));
    }
    double positive_part(double x) const {
        if (x <= -h) return 0.0;
        if (x >= h) return x;
        const double u = transition(x);
        const double u2 = u * u, u4 = u2 * u2;
        const double scaled = 2.0 * h * u4;
        const double poly = 2.5 + u * (-3.0 + u);
        return std::isnormal(scaled) ? scaled * poly : (2.0 * u4 * poly) * h;
    }
    double maximum(double l, double r) const {
        const double d = l - r;
        if (d >= h) return l;
        if (d <= -h) return r;
        return r + positive_part(d);
    }
    double minimum(double l, double r) const {
        const double d = l - r;
        if (d >= h) return r;
        if (d <= -h) return l;
        return l - positive_part(d);
    }
};

// fold_node (compile.rs:684-778): the value of an operation whose operands
// are all constants; `ok` false for a non-finite result.
double fold(const Source& s, const double* v, bool* ok) {
    double r = 0.0;
    switch (s.op) {
        case PE_GRAPH_LOG: r = std::log(v[0]); break;
        case PE_GRAPH_EXP: r = std::exp(v[0]); break;
        case PE_GRAPH_COMPARE: r = compare(s.cmp, v[0], v[1]) ? 1.0 : 0.0; break;
        case PE_GRAPH_ADD: r = v[0] + v[1]; break;
        case PE_GRAPH_SUBTRACT: r = v[0] - v[1]; break;
        case PE_GRAPH_MULTIPLY: r = v[0] * v[1]; break;
        case PE_GRAPH_DIVIDE: r = v[0] / v[1]; break;
        case PE_GRAPH_MINIMUM: r = v[0] <= v[1] ? v[0] : v[1]; break;
        case PE_GRAPH_MAXIMUM: r = v[0] >= v[1] ? v[0] : v[1]; break;
        case PE_GRAPH_INDICATOR: r = v[0] >= 0.0 ? 1.0 : 0.0; break;
        case PE_GRAPH_SMOOTH_MINIMUM: r = HostSmooth{s.value}.minimum(v[0], v[1]); break;
        case PE_GRAPH_SMOOTH_MAXIMUM: r = HostSmooth{s.value}.maximum(v[0], v[1]); break;
        case PE_GRAPH_SMOOTH_INDICATOR: r = HostSmooth{s.value}.indicator(v[0]); break;
        case PE_GRAPH_BARRIER_HIT_STATE: r = v[0] + (1.0 - v[0]) * v[1]; break;
        case PE_GRAPH_NEGATE: r = -v[0]; break;
        default: invalid("internal graph fold");
    }
    *ok = std::isfinite(r);
    return r;
}

uint16_t code_of(const Source& s) {
    switch (s.op) {
        case PE_GRAPH_TERMINAL_SPOT: return kGSpot;
        case PE_GRAPH_PRE_DIVIDEND_SPOT: return kGPreSpot;
        case PE_GRAPH_ADD: return kGAdd;
        case PE_GRAPH_SUBTRACT: return kGSub;
        case PE_GRAPH_MULTIPLY: return kGMul;
        case PE_GRAPH_DIVIDE: return kGDiv;
        case PE_GRAPH_MINIMUM: return kGMin;
        case PE_GRAPH_MAXIMUM: return kGMax;
        case PE_GRAPH_INDICATOR: return kGInd;
        case PE_GRAPH_NEGATE: return kGNeg;
        case PE_GRAPH_BARRIER_HIT_STATE: return kGHit;
        case PE_GRAPH_SMOOTH_MINIMUM: return kGSmoothMin;
        case PE_GRAPH_SMOOTH_MAXIMUM: return kGSmoothMax;
        case PE_GRAPH_SMOOTH_INDICATOR: return kGSmoothInd;
        case PE_GRAPH_LOG: return kGLog;
        case PE_GRAPH_EXP: return kGExp;
        case PE_GRAPH_SELECT: return kGSelect;
        case PE_GRAPH_COMPARE: return static_cast<uint16_t>(kGLess + static_cast<int>(s.cmp));
        default: invalid("internal graph opcode");
    }
}

}  // namespace

void compile_graph(const pe_product& pr, const std::function<double(int32_t)>& discount, GraphTape& g) {
    if (!pr.graph_nodes || pr.graph_node_count == 0) invalid("graph product without nodes");
    if (!pr.graph_cashflows || pr.graph_cashflow_count == 0) invalid("one payment date is required per graph output");
    if (pr.graph_cashflow_count > 64) invalid("graph outputs exceed the limit of 64");
    if (!pr.graph_underlyings || pr.graph_underlying_count == 0) invalid("graph underlyings must be nonempty and unique");
    std::set<uint32_t> declared(pr.graph_underlyings, pr.graph_underlyings + pr.graph_underlying_count);
    if (declared.size() != pr.graph_underlying_count) invalid("graph underlyings must be nonempty and unique");
    const uint32_t primary = pr.graph_underlyings[0];

    // source nodes by id (compile.rs:501-531)
    std::map<uint32_t, Source> nodes;
    bool branching = false;
    for (uint32_t i = 0; i < pr.graph_node_count; ++i) {
        const pe_graph_node& n = pr.graph_nodes[i];
        Source s{n.opcode, {n.operands[0], n.operands[1], n.operands[2]}, n.comparison, n.observation_day, n.value};
        const int k = operand_count(n.opcode);
        for (int j = k; j < 3; ++j) s.in[j] = 0;
        if (n.opcode == PE_GRAPH_LITERAL && !std::isfinite(n.value)) invalid("non-finite graph literal");
        if ((n.opcode == PE_GRAPH_SMOOTH_MINIMUM || n.opcode == PE_GRAPH_SMOOTH_MAXIMUM ||
             n.opcode == PE_GRAPH_SMOOTH_INDICATOR) &&
            !(std::isfinite(n.value) && n.value > 0.0))
            invalid("smoothing half width must be positive and finite");
        if (n.opcode == PE_GRAPH_COMPARE && (n.comparison < PE_GRAPH_LESS || n.comparison > PE_GRAPH_GREATER))
            invalid("unknown graph comparison");
        if (n.opcode == PE_GRAPH_TERMINAL_SPOT || n.opcode == PE_GRAPH_PRE_DIVIDEND_SPOT) {
            if (!declared.count(n.underlying_id)) invalid("graph contains an undeclared underlying");
            if (n.underlying_id != primary)
                unsupported("graph observation of an underlying other than the market's");
        }
        if (n.opcode == PE_GRAPH_SELECT) branching = true;
        if (!nodes.emplace(n.id, s).second) invalid("duplicate graph node id " + std::to_string(n.id));
    }
    for (uint32_t i = 0; i < pr.graph_cashflow_count; ++i)
        if (!nodes.count(pr.graph_cashflows[i].output_node)) invalid("unknown graph output node");
    for (const auto& [id, s] : nodes)
        for (int j = 0; j < operand_count(s.op); ++j)
            if (!nodes.count(s.in[j])) invalid("unknown operand of graph node " + std::to_string(id));

    // latest observation per node and the payment check (graph_product.rs:85-113);
    // a topological pass by recursion-free Kahn ordering
    std::map<uint32_t, int> indegree;
    std::map<uint32_t, std::vector<uint32_t>> outgoing;
    auto kahn = [&](std::vector<uint32_t>& order) {
        indegree.clear();
        outgoing.clear();
        for (const auto& [id, s] : nodes) {
            indegree[id] = operand_count(s.op);
            for (int j = 0; j < operand_count(s.op); ++j) outgoing[s.in[j]].push_back(id);
        }
        for (auto& [id, d] : outgoing) std::sort(d.begin(), d.end());
        std::set<uint32_t> ready;
        for (const auto& [id, d] : indegree)
            if (d == 0) ready.insert(id);
        order.clear();
        while (!ready.empty()) {
            const uint32_t id = *ready.begin();
            ready.erase(ready.begin());
            order.push_back(id);
            auto it = outgoing.find(id);
            if (it == outgoing.end()) continue;
            for (uint32_t dst : it->second)
                if (--indegree[dst] == 0) ready.insert(dst);
        }
        if (order.size() != nodes.size()) invalid("graph contains a cycle");
    };
    std::vector<uint32_t> order;
    kahn(order);
    std::map<uint32_t, int32_t> latest;   // INT32_MIN: no observation
    for (uint32_t id : order) {
        const Source& s = nodes[id];
        int32_t d = INT32_MIN;
        if (s.op == PE_GRAPH_TERMINAL_SPOT || s.op == PE_GRAPH_PRE_DIVIDEND_SPOT) d = s.day;
        for (int j = 0; j < operand_count(s.op); ++j) d = std::max(d, latest[s.in[j]]);
        latest[id] = d;
    }
    int32_t last_payment = INT32_MIN;
    for (uint32_t i = 0; i < pr.graph_cashflow_count; ++i) {
        const pe_graph_cashflow& c = pr.graph_cashflows[i];
        if (latest[c.output_node] != INT32_MIN && latest[c.output_node] > c.payment_day)
            invalid("cash flow depends on an observation after its payment");
        last_payment = std::max(last_payment, c.payment_day);
    }
    if (last_payment < 0) unsupported("graph cash flows paid before valuation");

    // one output: cash flows before the last payment carry DF(pay) / DF(last)
    // (compile_graph_cashflows); node ids continue after the largest
    uint32_t output = pr.graph_cashflows[0].output_node;
    if (pr.graph_cashflow_count > 1) {
        uint32_t next = nodes.rbegin()->first + 1;
        auto push = [&](Source s) {
            nodes[next] = s;
            return next++;
        };
        const double final_discount = discount(last_payment);
        bool have = false;
        uint32_t sum = 0;
        for (uint32_t i = 0; i < pr.graph_cashflow_count; ++i) {
            const pe_graph_cashflow& c = pr.graph_cashflows[i];
            uint32_t value = c.output_node;
            if (c.payment_day != last_payment) {
                const double ratio = discount(c.payment_day) / final_discount;
                if (!std::isfinite(ratio)) invalid("non-finite cash flow discount ratio");
                const uint32_t lit = push(Source{PE_GRAPH_LITERAL, {0, 0, 0}, PE_GRAPH_LESS, 0, ratio});
                value = push(Source{PE_GRAPH_MULTIPLY, {c.output_node, lit, 0}, PE_GRAPH_LESS, 0, 0.0});
            }
            sum = have ? push(Source{PE_GRAPH_ADD, {sum, value, 0}, PE_GRAPH_LESS, 0, 0.0}) : value;
            have = true;
        }
        output = sum;
        kahn(order);
    }

    // constant folding in topological order; a non-finite fold is an error
    // unless the graph branches (then it stays a run-time operation)
    std::map<uint32_t, bool> folded;
    std::map<uint32_t, double> folded_value;
    for (uint32_t id : order) {
        const Source& s = nodes[id];
        bool f = false;
        double v = 0.0;
        if (s.op == PE_GRAPH_LITERAL) {
            f = true;
            v = s.value;
        } else if (s.op == PE_GRAPH_SELECT) {
            if (folded[s.in[0]]) {
                const uint32_t pick = folded_value[s.in[0]] != 0.0 ? s.in[1] : s.in[2];
                if (folded[pick]) {
                    f = true;
                    v = folded_value[pick];
                }
            }
        } else if (s.op != PE_GRAPH_TERMINAL_SPOT && s.op != PE_GRAPH_PRE_DIVIDEND_SPOT) {
            const int k = operand_count(s.op);
            bool all = true;
            double in[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < k; ++j) {
                all = all && folded[s.in[j]];
                if (all) in[j] = folded_value[s.in[j]];
            }
            if (all) {
                bool ok = true;
                v = fold(s, in, &ok);
                if (ok) f = true;
                else if (!branching) invalid("non-finite constant fold at graph node " + std::to_string(id));
            }
        }
        folded[id] = f;
        folded_value[id] = v;
    }

    // the output's cone (reachable_nodes, compile.rs:669-682)
    std::set<uint32_t> reachable;
    std::vector<uint32_t> pending{output};
    while (!pending.empty()) {
        const uint32_t id = pending.back();
        pending.pop_back();
        if (!reachable.insert(id).second) continue;
        const Source& s = nodes[id];
        for (int j = 0; j < operand_count(s.op); ++j) pending.push_back(s.in[j]);
    }

    // observation days of the compiled tape (terminal_observations)
    std::set<int32_t> days, pre_days;
    for (uint32_t id : reachable) {
        const Source& s = nodes[id];
        if (folded[id]) continue;
        if (s.op == PE_GRAPH_TERMINAL_SPOT || s.op == PE_GRAPH_PRE_DIVIDEND_SPOT) {
            if (s.day < 0) unsupported("graph observation before valuation (no fixing history for graphs)");
            days.insert(s.day);
            if (s.op == PE_GRAPH_PRE_DIVIDEND_SPOT) pre_days.insert(s.day);
        }
    }
    g.days.assign(days.begin(), days.end());
    g.pre_days.assign(pre_days.begin(), pre_days.end());
    g.last_payment = last_payment;
    g.expiry_day = g.days.empty() ? last_payment : g.days.back();
    auto obs_of = [&](int32_t day) {
        return static_cast<int>(std::lower_bound(g.days.begin(), g.days.end(), day) - g.days.begin());
    };

    // schedule: an operation's stage is the latest observation index it
    // reads (-1: before the path); stable in topological order
    std::vector<uint32_t> live_order;
    for (uint32_t id : order)
        if (reachable.count(id) && !folded[id]) live_order.push_back(id);
    std::map<uint32_t, int> stage;
    for (uint32_t id : order) {
        if (!reachable.count(id)) continue;
        if (folded[id]) { stage[id] = -2; continue; }
        const Source& s = nodes[id];
        int st = -1;
        if (s.op == PE_GRAPH_TERMINAL_SPOT || s.op == PE_GRAPH_PRE_DIVIDEND_SPOT) st = obs_of(s.day);
        for (int j = 0; j < operand_count(s.op); ++j) st = std::max(st, stage[s.in[j]]);
        stage[id] = st;
    }
    std::stable_sort(live_order.begin(), live_order.end(),
                     [&](uint32_t x, uint32_t y) { return stage[x] < stage[y]; });

    // constants
    g.consts.clear();
    std::map<uint32_t, uint16_t> const_index;
    auto constant = [&](uint32_t id) -> uint16_t {
        auto it = const_index.find(id);
        if (it != const_index.end()) return it->second;
        if (g.consts.size() >= 0x3fffu) throw PlanError(PE_ERR_SIZE_LIMIT, "graph constants exceed 16383");
        const uint16_t k = static_cast<uint16_t>(kGraphConst | g.consts.size());
        g.consts.push_back(folded_value[id]);
        const_index[id] = k;
        return k;
    };

    // Fewer dispatched operations (the interpreter's cost is per operation):
    //  * a Multiply whose only consumer is the Add scheduled right after it
    //    becomes one a * b + c operation (two roundings in FP64 mode, which
    //    is built without contraction, so values are unchanged);
    //  * a spot whose every consumer runs in its own observation's segment is
    //    read from the observation instead of a register.
    // A fused product that is non-finite makes the sum non-finite, so the
    // poison rule is unchanged.
    std::map<uint32_t, std::vector<uint32_t>> consumers;
    for (uint32_t id : live_order) {
        const Source& s = nodes[id];
        for (int j = 0; j < operand_count(s.op); ++j)
            if (!folded[s.in[j]]) consumers[s.in[j]].push_back(id);
    }
    std::map<uint32_t, uint32_t> fused_into;   // Multiply -> its Add
    for (size_t pos = 0; pos + 1 < live_order.size(); ++pos) {
        const uint32_t m = live_order[pos], n = live_order[pos + 1];
        if (nodes[m].op != PE_GRAPH_MULTIPLY || nodes[n].op != PE_GRAPH_ADD || m == output) continue;
        const auto& cm = consumers[m];
        if (cm.size() != 1 || cm[0] != n) continue;                 // single use (also not Add(m, m))
        fused_into[m] = n;
    }
    auto emitter = [&](uint32_t id) {   // the operation that evaluates `id`
        auto it = fused_into.find(id);
        return it == fused_into.end() ? id : it->second;
    };
    std::set<uint32_t> inline_spot;
    for (uint32_t id : live_order) {
        const Source& s = nodes[id];
        if ((s.op != PE_GRAPH_TERMINAL_SPOT && s.op != PE_GRAPH_PRE_DIVIDEND_SPOT) || id == output) continue;
        bool same = !consumers[id].empty();
        for (uint32_t c : consumers[id]) same = same && stage[emitter(c)] == stage[id];
        if (same) inline_spot.insert(id);
    }

    // emitted operations: (node, code, operand nodes) in schedule order
    struct Emit {
        uint32_t id;
        uint16_t code;
        int n;
        uint32_t in[3];
    };
    std::vector<Emit> emit;
    for (uint32_t id : live_order) {
        if (fused_into.count(id) || inline_spot.count(id)) continue;
        const Source& s = nodes[id];
        Emit e{id, code_of(s), operand_count(s.op), {s.in[0], s.in[1], s.in[2]}};
        if (s.op == PE_GRAPH_ADD) {
            for (int j = 0; j < 2; ++j) {
                auto it = fused_into.find(s.in[j]);
                if (it == fused_into.end() || it->second != id) continue;
                const Source& m = nodes[s.in[j]];
                e = Emit{id, kGMulAdd, 3, {m.in[0], m.in[1], s.in[1 - j]}};
                break;
            }
        }
        emit.push_back(e);
    }
    std::map<uint32_t, size_t> last_use;
    for (size_t pos = 0; pos < emit.size(); ++pos)
        for (int j = 0; j < emit[pos].n; ++j) last_use[emit[pos].in[j]] = pos;
    last_use[output] = SIZE_MAX;

    // linear-scan register assignment over the emitted schedule
    std::map<uint32_t, uint16_t> reg;
    uint32_t free_mask = (kGraphMaxRegs == 32) ? 0xffffffffu : ((1u << kGraphMaxRegs) - 1u);
    uint32_t used = 0;
    g.ops.clear();
    g.segs.assign(g.days.size() + 2, 0u);
    int current_stage = -1;
    g.segs[0] = 0;
    auto operand = [&](uint32_t node) -> uint16_t {
        if (folded[node]) return constant(node);
        if (inline_spot.count(node))
            return nodes[node].op == PE_GRAPH_TERMINAL_SPOT ? kGraphPost : kGraphPre;
        return reg.at(node);
    };
    for (size_t pos = 0; pos < emit.size(); ++pos) {
        const Emit& e = emit[pos];
        const Source& s = nodes[e.id];
        while (current_stage < stage[e.id]) g.segs[static_cast<size_t>(++current_stage) + 1] = static_cast<uint32_t>(g.ops.size());
        GraphOp op{};
        op.code = e.code;
        op.param = s.value;
        uint16_t in[3] = {0, 0, 0};
        for (int j = 0; j < e.n; ++j) in[j] = operand(e.in[j]);
        op.a = in[0];
        op.b = in[1];
        op.c = in[2];
        double* kv[3] = {&op.ka, &op.kb, &op.kc};
        for (int j = 0; j < e.n; ++j)
            if (in[j] & kGraphConst) *kv[j] = g.consts[in[j] & (kGraphConst - 1u)];
        for (int j = 0; j < e.n; ++j)
            if (in[j] < kGraphMaxRegs && last_use[e.in[j]] == pos) free_mask |= 1u << in[j];
        if (free_mask == 0)
            throw PlanError(PE_ERR_SIZE_LIMIT, "graph needs more than " + std::to_string(kGraphMaxRegs) +
                                                   " live values at once");
        const uint16_t r = static_cast<uint16_t>(__builtin_ctz(free_mask));
        free_mask &= ~(1u << r);
        used = std::max<uint32_t>(used, r + 1u);
        op.out = r;
        reg[e.id] = r;
        g.ops.push_back(op);
    }
    while (current_stage < static_cast<int>(g.days.size())) g.segs[static_cast<size_t>(++current_stage) + 1] = static_cast<uint32_t>(g.ops.size());
    g.output = folded[output] ? constant(output) : reg.at(output);
    g.regs = static_cast<uint16_t>(std::max<uint32_t>(used, 1u));
}

}  // namespace pe
