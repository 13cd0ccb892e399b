// ORACLE -- test infrastructure only.  Restates the reference's host-side
// market inputs in FP64:
//   curves       crates/pricing/src/market/curve.rs:39-212
//   forward      crates/pricing/src/market/forward.rs:129-161
//   LV grid      crates/pricing/src/market/local_variance.rs:128-415, 417-478, 607-653
//   SSVI/Dupire  crates/pricing/src/market/ssvi.rs:61-186, 226-275, 340-475
//   theta PCHIP  crates/pricing/src/market/pchip.rs:27-238
//   LV grid time crates/pricing/src/engine/processes/local_vol.rs:166-237
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace oracle {

inline double rust_signum(double x) { return std::isnan(x) ? x : std::copysign(1.0, x); }

// Index of the last node ordered at or before `value`, clamped to the last
// cell: the binary_search_by(total_cmp) form of lower_cell (local_variance.rs:607-612).
inline size_t lower_cell(const std::vector<double>& nodes, double value) {
    // Finite inputs only (validated by the callers), so total_cmp == operator<
    // except -0.0 < +0.0, which cannot occur for the node sets used here.
    size_t lo = 0, hi = nodes.size();
    // number of nodes <= value
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (nodes[mid] <= value) lo = mid + 1; else hi = mid;
    }
    const size_t count = lo;
    const size_t idx = count == 0 ? 0 : count - 1;
    return std::min(idx, nodes.size() - 2);
}

// --------------------------------------------------------------- curves --
struct Curve {
    std::vector<double> times;
    std::vector<double> log_discounts;

    void init(const double* t, const double* df, uint32_t n) {
        if (n < 2) throw std::invalid_argument("curve needs two pillars");
        times.assign(t, t + n);
        if (times[0] != 0.0) throw std::invalid_argument("curve must start at t=0");
        if (df[0] != 1.0) throw std::invalid_argument("curve must start at DF=1");
        for (uint32_t i = 0; i < n; ++i) {
            if (!std::isfinite(t[i]) || t[i] < 0.0) throw std::invalid_argument("bad pillar time");
            if (!std::isfinite(df[i]) || df[i] <= 0.0) throw std::invalid_argument("bad discount factor");
            if (i > 0 && !(t[i] > t[i - 1])) throw std::invalid_argument("unsorted pillars");
        }
        times[0] = 0.0;
        log_discounts.resize(n);
        for (uint32_t i = 0; i < n; ++i) log_discounts[i] = std::log(df[i]);
    }

    double segment(size_t l, size_t r, double t) const {
        const double w = (t - times[l]) / (times[r] - times[l]);
        return log_discounts[l] + w * (log_discounts[r] - log_discounts[l]);
    }

    double log_discount(double t) const {
        if (!std::isfinite(t) || t < 0.0) throw std::invalid_argument("bad curve query time");
        if (t == 0.0) t = 0.0;
        // binary_search_by(total_cmp)
        auto it = std::lower_bound(times.begin(), times.end(), t);
        if (it != times.end() && *it == t) return log_discounts[static_cast<size_t>(it - times.begin())];
        const size_t right = static_cast<size_t>(it - times.begin());
        if (right < times.size()) return segment(right - 1, right, t);
        const size_t last = times.size() - 1;
        return segment(last - 1, last, t);
    }

    double discount(double t) const { return std::exp(log_discount(t)); }
};

struct Forward {  // forward.rs:129-161 (no discrete dividends)
    double spot;
    Curve discount_curve;
    Curve dividend_curve;

    double forward(double t) const {
        const double d = discount_curve.log_discount(t);
        const double q = dividend_curve.log_discount(t);
        const double f = spot * std::exp(q - d);
        if (!std::isfinite(f) || f <= 0.0) throw std::runtime_error("non-finite forward");
        return f;
    }
};

// ---------------------------------------------------- local variance grid --
struct Interp {
    double value;
    size_t t_cell, x_cell;
    double t_weight, x_weight;
    int boundary;  // 0 in range, 1 left flat, 2 right flat
};

struct LocalVarianceGrid {
    std::vector<double> t_nodes, x_nodes, values;
    double floor = 0.0, cap = 0.0;

    void init(const double* t, uint32_t nt, const double* x, uint32_t nx, const double* v, double fl,
              double cp) {
        if (nt < 2 || nx < 2) throw std::invalid_argument("LV grid needs 2x2 nodes");
        t_nodes.assign(t, t + nt);
        x_nodes.assign(x, x + nx);
        values.assign(v, v + static_cast<size_t>(nt) * nx);
        floor = fl;
        cap = cp;
        for (uint32_t i = 0; i < nt; ++i)
            if (!std::isfinite(t[i]) || t[i] < 0.0 || (i > 0 && !(t[i] > t[i - 1])))
                throw std::invalid_argument("bad LV time node");
        for (uint32_t i = 0; i < nx; ++i)
            if (!std::isfinite(x[i]) || (i > 0 && !(x[i] > x[i - 1])))
                throw std::invalid_argument("bad LV x node");
        for (double val : values)
            if (!std::isfinite(val) || val < floor || val > cap) throw std::invalid_argument("bad LV value");
    }

    // interpolate (local_variance.rs:273-307) + interpolate_in_cells (:340-385)
    Interp interpolate(double time, double x_in) const {
        if (!std::isfinite(time) || time < t_nodes.front() || time > t_nodes.back())
            throw std::runtime_error("LV time query outside grid");
        if (!std::isfinite(x_in)) throw std::runtime_error("non-finite log-moneyness");
        const size_t tc = lower_cell(t_nodes, time);
        double x = x_in;
        int boundary = 0;
        if (x_in < x_nodes.front()) { x = x_nodes.front(); boundary = 1; }
        else if (x_in > x_nodes.back()) { x = x_nodes.back(); boundary = 2; }
        const size_t xc = lower_cell(x_nodes, x);
        const double tw = (time - t_nodes[tc]) / (t_nodes[tc + 1] - t_nodes[tc]);
        const double xw = (x - x_nodes[xc]) / (x_nodes[xc + 1] - x_nodes[xc]);
        const size_t nx = x_nodes.size();
        const size_t row = tc * nx, next = row + nx;
        const double v00 = values[row + xc], v01 = values[row + xc + 1];
        const double v10 = values[next + xc], v11 = values[next + xc + 1];
        const double lower = v00 * (1.0 - xw) + v01 * xw;
        const double upper = v10 * (1.0 - xw) + v11 * xw;
        return Interp{lower * (1.0 - tw) + upper * tw, tc, xc, tw, xw, boundary};
    }
};

// LocalVolTimeGrid::compile (local_vol.rs:166-237)
inline size_t interval_step_count(double interval, double maximum_step) {
    const double raw = std::ceil(interval / maximum_step);
    if (!std::isfinite(raw) || raw > 1e15) throw std::invalid_argument("LV step count overflow");
    return std::max<size_t>(1, static_cast<size_t>(raw));
}

struct TimeGrid {
    std::vector<double> nodes;
    std::vector<size_t> event_node_indices;

    void compile(std::vector<double> events, double maximum_step) {
        if (!std::isfinite(maximum_step) || maximum_step <= 0.0)
            throw std::invalid_argument("non-positive LV maximum step");
        if (events.empty()) throw std::invalid_argument("empty LV event grid");
        for (double e : events)
            if (!std::isfinite(e) || e < 0.0 || (e == 0.0 && std::signbit(e)))
                throw std::invalid_argument("invalid LV event time");
        std::sort(events.begin(), events.end());
        events.erase(std::unique(events.begin(), events.end()), events.end());
        if (events[0] != 0.0) events.insert(events.begin(), 0.0);
        if (events.back() <= 0.0) throw std::invalid_argument("non-positive final LV time");
        nodes.clear();
        event_node_indices.clear();
        nodes.push_back(0.0);
        event_node_indices.push_back(0);
        for (size_t i = 0; i + 1 < events.size(); ++i) {
            const double left = events[i], right = events[i + 1];
            if (right <= left) continue;
            const double interval = right - left;
            const size_t steps = interval_step_count(interval, maximum_step);
            const double substep = interval / static_cast<double>(steps);
            for (size_t k = 1; k < steps; ++k) nodes.push_back(left + substep * static_cast<double>(k));
            nodes.push_back(right);
            event_node_indices.push_back(nodes.size() - 1);
        }
    }
    size_t steps() const { return nodes.size() - 1; }
};

// ------------------------------------------------------------ SSVI/Dupire --
struct ThetaPchip {  // pchip.rs
    std::vector<double> times, values, derivatives;
    double terminal_slope = 0.0;

    static double endpoint_derivative(double aw, double nw, double as, double ns) {
        double d = ((2.0 * aw + nw) * as - aw * ns) / (aw + nw);
        if (rust_signum(d) != rust_signum(as)) {
            d = 0.0;
        } else if (rust_signum(as) != rust_signum(ns) && std::fabs(d) > 3.0 * std::fabs(as)) {
            d = 3.0 * as;
        }
        return d;
    }

    void init(const double* t, const double* v, uint32_t n, double slope) {
        if (n < 2) throw std::invalid_argument("theta curve needs two knots");
        times.assign(t, t + n);
        values.assign(v, v + n);
        terminal_slope = slope;
        std::vector<double> widths(n - 1), secants(n - 1);
        for (uint32_t i = 0; i + 1 < n; ++i) {
            widths[i] = t[i + 1] - t[i];
            secants[i] = (v[i + 1] - v[i]) / widths[i];
        }
        derivatives.assign(n, 0.0);
        if (n == 2) {
            derivatives[0] = secants[0];
            derivatives[1] = secants[0];
            return;
        }
        derivatives[0] = endpoint_derivative(widths[0], widths[1], secants[0], secants[1]);
        for (uint32_t i = 1; i + 1 < n; ++i) {
            const double l = secants[i - 1], r = secants[i];
            if (l == 0.0 || r == 0.0 || rust_signum(l) != rust_signum(r)) {
                derivatives[i] = 0.0;
            } else {
                const double lw = widths[i - 1], rw = widths[i];
                const double w1 = 2.0 * rw + lw;
                const double w2 = rw + 2.0 * lw;
                derivatives[i] = (w1 + w2) / (w1 / l + w2 / r);
            }
        }
        const size_t last = n - 2;
        derivatives[n - 1] = endpoint_derivative(widths[last], widths[last - 1], secants[last], secants[last - 1]);
    }

    void evaluate(double time, double* theta, double* derivative) const {
        if (time == 0.0) time = 0.0;
        if (time < times[0]) {
            const double d = values[0] / times[0];
            *theta = time * d;
            *derivative = d;
            return;
        }
        auto it = std::lower_bound(times.begin(), times.end(), time);
        const size_t right = static_cast<size_t>(it - times.begin());
        if (it != times.end() && *it == time) {
            *theta = values[right];
            *derivative = derivatives[right];
            return;
        }
        if (right < times.size()) {
            const size_t left = right - 1;
            const double width = times[right] - times[left];
            const double s = (time - times[left]) / width;
            const double s2 = s * s;
            const double s3 = s2 * s;
            const double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
            const double h10 = s3 - 2.0 * s2 + s;
            const double h01 = -2.0 * s3 + 3.0 * s2;
            const double h11 = s3 - s2;
            *theta = h00 * values[left] + h10 * width * derivatives[left] + h01 * values[right] +
                     h11 * width * derivatives[right];
            const double d00 = (6.0 * s2 - 6.0 * s) / width;
            const double d10 = 3.0 * s2 - 4.0 * s + 1.0;
            const double d01 = (-6.0 * s2 + 6.0 * s) / width;
            const double d11 = 3.0 * s2 - 2.0 * s;
            *derivative = d00 * values[left] + d10 * derivatives[left] + d01 * values[right] +
                          d11 * derivatives[right];
            return;
        }
        const size_t last = times.size() - 1;
        *theta = values[last] + terminal_slope * (time - times[last]);
        *derivative = terminal_slope;
    }
};

struct StandardSsvi {  // ssvi.rs:340-475 with PhiSpec::PowerLaw (:123-136)
    ThetaPchip theta_curve;
    double rho, eta, gamma;

    // Returns (w, dw/dk, d2w/dk2, dw/dt).
    void total_variance_derivatives(double time, double k, double out[4]) const {
        if (!std::isfinite(time) || time <= 0.0) throw std::invalid_argument("SSVI time must be positive");
        double theta, theta_d;
        theta_curve.evaluate(time, &theta, &theta_d);
        const double phi = eta / (std::pow(theta, gamma) * std::pow(1.0 + theta, 1.0 - gamma));
        const double phi_d = -phi * (gamma / theta + (1.0 - gamma) / (1.0 + theta));
        const double y = phi * k;
        const double a = y + rho;
        const double q = std::sqrt(a * a + 1.0 - rho * rho);
        const double shape = 1.0 + rho * y + q;
        const double w = theta * shape / 2.0;
        const double common = rho + a / q;
        const double wk = theta * phi * common / 2.0;
        const double wkk = theta * phi * phi * (1.0 - rho * rho) / (2.0 * q * q * q);
        const double theta_partial = shape / 2.0 + theta * k * phi_d * common / 2.0;
        const double wt = theta_partial * theta_d;
        out[0] = w; out[1] = wk; out[2] = wkk; out[3] = wt;
    }
};

// durrleman_density_factor (ssvi.rs:226-275)
inline double durrleman_density_factor(double k, double w, double wk, double wkk) {
    if (!std::isfinite(w) || w <= 0.0) throw std::runtime_error("non-positive total variance");
    const double u = 1.0 - k * wk / (2.0 * w);
    const double g = u * u - (wk * wk / 4.0) * (1.0 / w + 1.0 / 4.0) + wkk / 2.0;
    if (!std::isfinite(g) || g <= 0.0) throw std::runtime_error("non-positive density factor");
    return g;
}

// LocalVarianceGrid::from_surface (local_variance.rs:149-206) with
// classify_local_variance (:583-605).
inline std::vector<double> dupire_from_ssvi(const StandardSsvi& s, const std::vector<double>& t_nodes,
                                            const std::vector<double>& x_nodes, double floor, double cap,
                                            uint32_t* repairs) {
    // validate_nodes / validate_floor_cap (local_variance.rs:510-565)
    auto validate = [](const std::vector<double>& v, bool reject_negative) {
        if (v.size() < 2) throw std::invalid_argument("local variance grid needs two nodes per axis");
        for (size_t i = 0; i < v.size(); ++i) {
            if (!std::isfinite(v[i]) || (reject_negative && v[i] < 0.0)) throw std::invalid_argument("bad node");
            if (i > 0 && !(v[i] > v[i - 1])) throw std::invalid_argument("unsorted nodes");
        }
    };
    validate(t_nodes, true);
    validate(x_nodes, false);
    if (!std::isfinite(floor) || floor <= 0.0 || !std::isfinite(cap) || cap < floor)
        throw std::invalid_argument("bad floor/cap");
    std::vector<double> out;
    out.reserve(t_nodes.size() * x_nodes.size());
    uint32_t count = 0;
    for (size_t i = 0; i < t_nodes.size(); ++i) {
        const double st = t_nodes[i] == 0.0 ? t_nodes[1] : t_nodes[i];
        for (size_t j = 0; j < x_nodes.size(); ++j) {
            double v[4];
            s.total_variance_derivatives(st, x_nodes[j], v);
            if (!std::isfinite(v[3])) throw std::runtime_error("non-finite time derivative");
            const double g = durrleman_density_factor(x_nodes[j], v[0], v[1], v[2]);
            const double raw = v[3] / g;
            double value = raw;
            bool repaired = true;
            if (std::isnan(raw)) value = floor;
            else if (raw == -INFINITY) value = floor;
            else if (raw == INFINITY) value = cap;
            else if (raw <= 0.0) value = floor;
            else if (raw < floor) value = floor;
            else if (raw > cap) value = cap;
            else repaired = false;
            if (repaired) ++count;
            out.push_back(value);
        }
    }
    if (repairs) *repairs = count;
    return out;
}

// piecewise_sinh_log_moneyness_nodes (local_variance.rs:417-478)
inline double sinh_ratio(double shape, double fraction) {
    return shape == 0.0 ? fraction : std::sinh(shape * fraction) / std::sinh(shape);
}

inline std::vector<double> piecewise_sinh_nodes(double left, double right, size_t nl, size_t nr,
                                                double left_shape, double right_shape) {
    std::vector<double> nodes;
    for (size_t i = 0; i <= nl; ++i) {
        const double fraction = static_cast<double>(i) / static_cast<double>(nl);
        const double ratio = sinh_ratio(left_shape, 1.0 - fraction);
        nodes.push_back(i == nl ? 0.0 : left * ratio);
    }
    for (size_t i = 1; i <= nr; ++i) {
        const double fraction = static_cast<double>(i) / static_cast<double>(nr);
        nodes.push_back(right * sinh_ratio(right_shape, fraction));
    }
    return nodes;
}

}  // namespace oracle
