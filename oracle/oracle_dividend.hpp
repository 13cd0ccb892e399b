// ORACLE -- test infrastructure only.  Restates the reference's escrowed
// affine discrete-dividend model in FP64:
//   crates/pricing/src/market/dividend.rs:96-211 (quotes, compiled events,
//   AffineDividendCoordinate), :277-520 (AffineDividendTransform: reserve,
//   carry, coordinate_after_time, event_timeline), :622-668 (compile/validate)
//   crates/pricing/src/market/forward.rs:49-70, 129-161 (EquityForward)
// Spot reconstruction is S = A*S0 + B*F with the canonical process F.
#pragma once

#include <cmath>
#include <stdexcept>
#include <vector>

#include "oracle_market.hpp"

namespace oracle {

struct Coord {  // AffineDividendCoordinate
    double a = 0.0, b = 1.0, spot_scale = 1.0;

    // at_spot (dividend.rs:163-173): refit the spot-dependent escrow scale.
    Coord at_spot(double original, double spot) const {
        if (original == spot) return *this;
        Coord c;
        c.a = a * original / spot;
        c.b = spot_scale - (spot_scale - b) * original / spot;
        c.spot_scale = spot_scale;
        return c;
    }
};

struct DividendEventC {  // CompiledDividendEvent
    double ex_time, fixed_cash, alpha, beta;
};

struct Dividends {  // AffineDividendTransform with carry curves
    double spot = 0.0;
    std::vector<DividendEventC> events;
    const Curve* discount = nullptr;
    const Curve* dividend = nullptr;
    double reserve = 0.0;

    bool empty() const { return events.empty(); }

    void init(double s0, const double* ex_time, const double* cash, const double* beta, size_t n, const Curve* p,
              const Curve* q) {
        spot = s0;
        discount = p;
        dividend = q;
        events.clear();
        for (size_t i = 0; i < n; ++i) {  // validate_events
            if (!std::isfinite(ex_time[i]) || ex_time[i] < 0.0) throw std::invalid_argument("bad dividend time");
            if (!(std::isfinite(cash[i]) && cash[i] >= 0.0)) throw std::invalid_argument("bad dividend cash");
            if (!(std::isfinite(beta[i]) && beta[i] >= 0.0 && beta[i] < 1.0))
                throw std::invalid_argument("bad dividend beta");
            if (i > 0 && ex_time[i] <= ex_time[i - 1]) throw std::invalid_argument("unsorted dividend events");
            const double alpha = cash[i] / s0;  // compile_event
            if (!std::isfinite(alpha)) throw std::invalid_argument("non-finite dividend alpha");
            events.push_back({ex_time[i], cash[i], alpha, beta[i]});
        }
        // compile_reserve (dividend.rs:340-361)
        double proportional = 1.0, r = 0.0;
        for (const auto& e : events) {
            proportional *= 1.0 - e.beta;
            const double d = p->discount(e.ex_time) / q->discount(e.ex_time);
            r += e.fixed_cash / proportional * d;
        }
        const double residual = spot - r;
        if (!std::isfinite(r) || !std::isfinite(residual) || residual <= 0.0)
            throw std::invalid_argument("non-positive escrowed spot");
        reserve = r;
    }

    Coord initial() const {
        Coord c;
        c.a = reserve / spot;
        c.b = (spot - reserve) / spot;
        c.spot_scale = 1.0;
        return c;
    }

    Coord carry(Coord c, double start, double end) const {  // dividend.rs:373-398
        if (c.a == 0.0 || start == end) return c;
        const double g = dividend->log_discount(end) - discount->log_discount(end) - dividend->log_discount(start) +
                         discount->log_discount(start);
        c.a *= std::exp(g);
        if (!std::isfinite(c.a)) throw std::runtime_error("non-finite carried affine A");
        return c;
    }

    static Coord after(Coord c, const DividendEventC& e) {  // dividend.rs:181-205
        Coord n;
        n.a = (1.0 - e.beta) * c.a - e.alpha;
        n.b = (1.0 - e.beta) * c.b;
        n.spot_scale = (1.0 - e.beta) * c.spot_scale;
        if (!std::isfinite(n.a) || !std::isfinite(n.b)) throw std::runtime_error("non-finite dividend transform");
        return n;
    }

    Coord after_time(double t) const {  // coordinate_after_time (dividend.rs:470-497)
        if (!std::isfinite(t) || t < 0.0) throw std::invalid_argument("bad dividend query time");
        Coord c = initial();
        double prev = 0.0;
        for (const auto& e : events) {
            if (!(e.ex_time <= t)) break;
            c = carry(c, prev, e.ex_time);
            c = after(c, e);
            prev = e.ex_time;
        }
        c = carry(c, prev, t);
        bool all = true;
        for (const auto& e : events)
            if (!(e.ex_time <= t || e.fixed_cash == 0.0)) all = false;
        if (all) c.a = 0.0;
        return c;
    }

    struct Entry {
        double ex_time;
        Coord before, after;
    };

    std::vector<Entry> timeline() const {  // event_timeline (dividend.rs:524-547)
        std::vector<Entry> out;
        Coord c = initial();
        double prev = 0.0;
        long last_cash = -1;
        for (size_t i = 0; i < events.size(); ++i)
            if (events[i].fixed_cash != 0.0) last_cash = static_cast<long>(i);
        for (size_t i = 0; i < events.size(); ++i) {
            const Coord before = carry(c, prev, events[i].ex_time);
            Coord aft = after(before, events[i]);
            if (last_cash < 0 || static_cast<long>(i) >= last_cash) aft.a = 0.0;
            out.push_back({events[i].ex_time, before, aft});
            c = aft;
            prev = events[i].ex_time;
        }
        return out;
    }
};

// EquityForward::evaluate with its affine coordinate (forward.rs:129-161).
struct ForwardEval {
    double forward;
    Coord coord;
};

inline ForwardEval evaluate_forward(const Forward& fwd, const Dividends* div, double t) {
    ForwardEval e;
    e.forward = fwd.forward(t);
    e.coord = (div && !div->empty()) ? div->after_time(t) : Coord{};
    const double spot_contract = e.coord.a * fwd.spot + e.coord.b * e.forward;
    if (!std::isfinite(spot_contract) || spot_contract <= 0.0) throw std::runtime_error("non-finite forward");
    return e;
}

}  // namespace oracle
