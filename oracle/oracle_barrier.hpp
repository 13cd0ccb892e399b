// ORACLE -- test infrastructure only.  Restates the reference's continuous
// barrier monitoring by Brownian-bridge survival in log space, FP64:
//   crates/pricing/src/engine/payoff/barrier_bridge.rs:101-122 (transformed barrier),
//     :152-271 (smoothed endpoint), :273-426 (smoothed interval),
//     :446-595 (exact interval), :597-654 (exact path), :656-683 (helpers)
//   crates/pricing/src/engine/payoff/barrier.rs:28-40 (resolved survival),
//     :57-146 (smoothed path, reverse, diagnostics), :214-231 (exact
//     diagnostics), :260-320 (hit factors), :322-372 (payoff terms)
#pragma once

#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

#include "oracle_dividend.hpp"
#include "oracle_payoff.hpp"

namespace oracle {

constexpr double kLogMaxExpArgument = 6.564958884017094;   // barrier_bridge.rs:9
constexpr double kLogMinPositive = -744.4400719213812;      // barrier_bridge.rs:10
constexpr double kLn2 = 0.693147180559945309417232121458176568;

enum class BridgeStatus { FiniteCorrection, TouchedEndpoint, ZeroVariance, SurvivalUnderflow, CertainSurvival };

inline void bridge_positive(double v, const char* what) {
    if (!(std::isfinite(v) && v > 0.0)) throw std::runtime_error(std::string("barrier bridge ") + what + " must be positive");
}
inline void bridge_variance(double v) {
    if (!(std::isfinite(v) && v >= 0.0)) throw std::runtime_error("barrier bridge local variance must be non-negative");
}

inline double transformed_barrier(double spot_barrier, double initial_spot, const Coord& c) {
    bridge_positive(spot_barrier, "spot barrier");
    bridge_positive(initial_spot, "initial spot");
    if (!std::isfinite(c.b) || c.b <= 0.0) throw std::runtime_error("barrier bridge affine scale must be positive");
    const double numerator = spot_barrier - c.a * initial_spot;
    const double barrier = numerator / c.b;
    if (!std::isfinite(barrier) || barrier <= 0.0) throw std::runtime_error("transformed barrier must be positive");
    return barrier;
}

// Shared log-survival core of the exact and smoothed intervals.
struct LogSurvival {
    BridgeStatus status;
    double log_survival;
    double led;  // d(log survival)/d(log exponent); 0 for deterministic statuses
    bool deterministic;
};

inline LogSurvival bridge_log_survival(double ld, double rd, double iv) {
    const double log_exponent = kLn2 + std::log(ld) + std::log(rd) - std::log(iv);
    if (log_exponent >= kLogMaxExpArgument) return {BridgeStatus::CertainSurvival, 0.0, 0.0, true};
    if (log_exponent <= kLogMinPositive) return {BridgeStatus::SurvivalUnderflow, log_exponent, 1.0, false};
    const double exponent = std::exp(log_exponent);
    const double survival = -std::expm1(-exponent);
    return {BridgeStatus::FiniteCorrection, std::log(survival), exponent / std::expm1(exponent), false};
}

// ------------------------------------------------------------ exact path --
struct ExactInterval {
    BridgeStatus status;
    double log_survival;
    double d_left_state = 0, d_right_state = 0, d_left_barrier = 0, d_right_barrier = 0;
    double d_left_var = 0, d_right_var = 0, d_dt = 0;
};

inline ExactInterval exact_interval(bool up, double ls, double rs, double lb, double rb, double lv, double rv,
                                    double dt) {
    bridge_positive(ls, "left state");
    bridge_positive(rs, "right state");
    bridge_positive(lb, "left transformed barrier");
    bridge_positive(rb, "right transformed barrier");
    bridge_variance(lv);
    bridge_variance(rv);
    if (!std::isfinite(dt) || dt <= 0.0) throw std::runtime_error("barrier bridge interval length must be positive");
    const double ld = up ? std::log(lb) - std::log(ls) : std::log(ls) - std::log(lb);
    const double rd = up ? std::log(rb) - std::log(rs) : std::log(rs) - std::log(rb);
    ExactInterval out{};
    if (ld <= 0.0 || rd <= 0.0) {
        out.status = BridgeStatus::TouchedEndpoint;
        out.log_survival = -std::numeric_limits<double>::infinity();
        return out;
    }
    const double avg = 0.5 * lv + 0.5 * rv;
    const double iv = avg * dt;
    if (!std::isfinite(iv)) throw std::runtime_error("non-finite integrated variance");
    if (iv == 0.0) {
        out.status = BridgeStatus::ZeroVariance;
        out.log_survival = 0.0;
        return out;
    }
    const LogSurvival s = bridge_log_survival(ld, rd, iv);
    out.status = s.status;
    out.log_survival = s.log_survival;
    if (s.deterministic) return out;
    const double ldd = s.led / ld, rdd = s.led / rd;
    const double ivd = -s.led / iv;
    const double sign = up ? 1.0 : -1.0;
    const double avd = ivd * dt;
    out.d_left_state = -sign * ldd / ls;
    out.d_right_state = -sign * rdd / rs;
    out.d_left_barrier = sign * ldd / lb;
    out.d_right_barrier = sign * rdd / rb;
    out.d_left_var = 0.5 * avd;
    out.d_right_var = 0.5 * avd;
    out.d_dt = ivd * avg;
    return out;
}

struct BridgeInput {
    double ls, rs, lb, rb, lv, rv, dt;
};

struct BridgeCounts {
    uint32_t interval = 0, finite = 0, touched = 0, zero = 0, underflow = 0, certain = 0;
    void add(BridgeStatus s) {
        switch (s) {
            case BridgeStatus::FiniteCorrection: ++finite; break;
            case BridgeStatus::TouchedEndpoint: ++touched; break;
            case BridgeStatus::ZeroVariance: ++zero; break;
            case BridgeStatus::SurvivalUnderflow: ++underflow; break;
            case BridgeStatus::CertainSurvival: ++certain; break;
        }
    }
};

struct ExactPath {  // BarrierBridgePath
    std::vector<ExactInterval> intervals;
    double log_survival = 0.0;
    BridgeCounts counts;

    void evaluate(bool up, const std::vector<BridgeInput>& in) {
        intervals.clear();
        log_survival = 0.0;
        counts = BridgeCounts{};
        counts.interval = static_cast<uint32_t>(in.size());
        for (const auto& x : in) {
            const ExactInterval it = exact_interval(up, x.ls, x.rs, x.lb, x.rb, x.lv, x.rv, x.dt);
            counts.add(it.status);
            log_survival += it.log_survival;
            intervals.push_back(it);
        }
    }
    double survival() const { return std::exp(log_survival); }
};

// -------------------------------------------------------- smoothed path --
struct SEndpoint {  // SmoothedBarrierBridgeEndpoint
    double hit = 0, safety = 1, eld = 0;
    double h_state = 0, h_barrier = 0, h_scale = 0;   // hit-weight derivatives
    double e_state = 0, e_barrier = 0, e_scale = 0;   // effective-log-distance derivatives

    // reverse(hit_adjoint, eld_adjoint).state
    double reverse_state(double ha, double ea) const { return ha * h_state + ea * e_state; }
};

inline SEndpoint smoothed_endpoint(bool up, double state, double tb, double scale, const Smooth& sm) {
    bridge_positive(state, "state");
    bridge_positive(tb, "transformed barrier");
    bridge_positive(scale, "affine scale");
    const double sign = up ? 1.0 : -1.0;
    const double diff = state - tb;
    const double shd = scale * sign * diff;
    SEndpoint e;
    double hit_first;
    sm.indicator_with_complement(shd, &e.hit, &hit_first, &e.safety);
    double safe_value, safe_first, safe_comp;
    sm.positive_part(-shd, &safe_value, &safe_first, &safe_comp);
    const double tsb = scale * tb;
    const double denominator = up ? tsb - safe_value : tsb + safe_value;
    if (!std::isfinite(denominator) || denominator <= 0.0)
        throw std::runtime_error("smoothed barrier safe distance exceeds the transformed barrier");
    e.eld = up ? -std::log1p(-safe_value / tsb) : std::log1p(safe_value / tsb);
    const double hit_x = hit_first;
    const double safe_x = -safe_first;
    const double distance_safe = 1.0 / denominator;
    const double dbp = -safe_value / (tsb * denominator);
    const double distance_x = distance_safe * safe_x;
    const double x_state = scale * sign;
    const double x_barrier = -x_state;
    const double x_scale = sign * diff;
    const double c_barrier = scale;
    const double c_scale = tb;
    e.h_state = hit_x * x_state;
    e.h_barrier = hit_x * x_barrier;
    e.h_scale = hit_x * x_scale;
    e.e_state = distance_x * x_state;
    e.e_barrier = distance_x * x_barrier + dbp * c_barrier;
    e.e_scale = distance_x * x_scale + dbp * c_scale;
    return e;
}

struct SInterval {  // SmoothedBarrierBridgeInterval
    BridgeStatus status;
    double log_survival;
    SEndpoint left, right;
    double d_ld = 0, d_rd = 0, d_lv = 0, d_rv = 0, d_dt = 0;
};

inline SInterval smoothed_interval(bool up, double ls, double lb, double lscale, double rs, double rb, double rscale,
                                   double lv, double rv, double dt, const Smooth& sm) {
    SInterval out{};
    out.left = smoothed_endpoint(up, ls, lb, lscale, sm);
    out.right = smoothed_endpoint(up, rs, rb, rscale, sm);
    bridge_variance(lv);
    bridge_variance(rv);
    if (!std::isfinite(dt) || dt <= 0.0) throw std::runtime_error("barrier bridge interval length must be positive");
    const double ld = out.left.eld, rd = out.right.eld;
    if (ld == 0.0 || rd == 0.0) {
        out.status = BridgeStatus::TouchedEndpoint;
        out.log_survival = -std::numeric_limits<double>::infinity();
        return out;
    }
    const double avg = 0.5 * lv + 0.5 * rv;
    const double iv = avg * dt;
    if (!std::isfinite(iv)) throw std::runtime_error("non-finite integrated variance");
    if (iv == 0.0) {
        out.status = BridgeStatus::ZeroVariance;
        out.log_survival = 0.0;
        return out;
    }
    const LogSurvival s = bridge_log_survival(ld, rd, iv);
    out.status = s.status;
    out.log_survival = s.log_survival;
    if (s.deterministic) return out;
    const double ivd = -s.led / iv;
    const double avd = ivd * dt;
    out.d_ld = s.led / ld;
    out.d_rd = s.led / rd;
    out.d_lv = 0.5 * avd;
    out.d_rv = 0.5 * avd;
    out.d_dt = ivd * avg;
    return out;
}

enum class HitKind { Endpoint, DividendJump };

struct HitFactor {  // SmoothedBarrierHitFactor
    HitKind kind;
    long state_index;  // -1: the initial state
    double safety;
    double state_derivative;
};

inline HitFactor endpoint_hit_factor(const SEndpoint& e, long index) {
    return {HitKind::Endpoint, index, e.safety, e.reverse_state(1.0, 0.0)};
}

inline HitFactor jump_hit_factor(bool up, double pre_spot, double post_spot, double pre_scale, double post_scale,
                                 double barrier, const Smooth& sm, long index) {
    const double sign = up ? 1.0 : -1.0;
    const double pre_d = sign * (pre_spot - barrier);
    const double post_d = sign * (post_spot - barrier);
    double score, lf, rf;
    sm.maximum(pre_d, post_d, &score, &lf, &rf);
    double hit, hit_first, safety;
    sm.indicator_with_complement(score, &hit, &hit_first, &safety);
    const double ssd = sign * (lf * pre_scale + rf * post_scale);
    return {HitKind::DividendJump, index, safety, hit_first * ssd};
}

struct SmoothedPath {  // SmoothedContinuousBarrierPath
    std::vector<SInterval> intervals;
    std::vector<HitFactor> hits;
    double log_bridge = 0.0, log_hit = 0.0;
    BridgeCounts counts;

    void finish() {
        log_bridge = 0.0;
        counts = BridgeCounts{};
        for (const auto& it : intervals) {
            log_bridge += it.log_survival;
            if (it.status != BridgeStatus::TouchedEndpoint) counts.add(it.status);
        }
        counts.interval = static_cast<uint32_t>(intervals.size());
        log_hit = 0.0;
        for (const auto& h : hits) log_hit = log_hit + std::log(h.safety);
    }
    double survival() const { return std::exp(log_bridge + log_hit); }
};

// ------------------------------------------------------------ payoff terms --
struct ContinuousBarrier {  // ContinuousBarrierRuntime
    double strike = 0, barrier = 0, notional = 1, rebate = 0;
    bool call = true, up = true, knock_in = false;
    double monitoring_end = 0;
    int historical_hit = -1;
    std::vector<size_t> bridge_obs;
    size_t expiry_obs = 0;

    bool resolved(double* s) const {
        if (historical_hit == 1) { *s = 0.0; return true; }
        if (monitoring_end < 0.0) { *s = 1.0; return true; }
        return false;
    }
};

struct BarrierTerms {
    double value, terminal_derivative, survival_derivative;
};

inline BarrierTerms continuous_barrier_payoff_terms(const ContinuousBarrier& b, double terminal, double survival) {
    double fixed;
    const bool is_resolved = b.resolved(&fixed);
    if (is_resolved) {
        const bool active = b.knock_in ? fixed == 0.0 : fixed == 1.0;
        if (!active) return {b.rebate, 0.0, 0.0};
    }
    const double signed_intrinsic = b.call ? terminal - b.strike : b.strike - terminal;
    const double terminal_sign = b.call ? 1.0 : -1.0;
    const double vanilla = (signed_intrinsic >= 0.0 ? signed_intrinsic : 0.0) * b.notional;
    const double vanilla_derivative = signed_intrinsic >= 0.0 ? terminal_sign * b.notional : 0.0;
    if (is_resolved) return {vanilla, vanilla_derivative, 0.0};
    if (!b.knock_in)
        return {b.rebate + survival * (vanilla - b.rebate), survival * vanilla_derivative, vanilla - b.rebate};
    return {vanilla + survival * (b.rebate - vanilla), (1.0 - survival) * vanilla_derivative, b.rebate - vanilla};
}

}  // namespace oracle
