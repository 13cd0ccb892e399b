// ORACLE -- test infrastructure only.  Restates the reference's least-squares
// Monte Carlo exercise policy in FP64:
//   PolynomialBasisSpec          crates/pricing/src/engine/mc/lsm.rs:328-496
//   FeatureScaling::fit          :537-575, standardize :597-617
//   prediction, decisions        :776-819
//   value_exercise_policy        :1347-1482
//   train_exercise_policy        :1486-1695
//   fit_exercise_decision        :1928-2065
//   fit_polynomial_regression    :2067-2220, standardize_features :2222-2245,
//                                evaluate_monomial :2247-2261
//   fit_cpqr                     :2374-2568, column norms :2594-2649
//   should_exercise              :2313-2320
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "oracle_numerics.hpp"

namespace oracle {

struct Basis {  // PolynomialBasisSpec
    uint32_t feature_count = 1;
    uint32_t max_degree = 3;
    std::vector<std::vector<uint32_t>> exponents;

    static void enumerate_degree(std::vector<std::vector<uint32_t>>& out, std::vector<uint32_t>& cur, size_t f,
                                 uint32_t remaining) {
        if (f + 1 == cur.size()) {
            cur[f] = remaining;
            out.push_back(cur);
            return;
        }
        for (uint32_t e = remaining + 1; e-- > 0;) {
            cur[f] = e;
            enumerate_degree(out, cur, f + 1, remaining - e);
        }
    }

    void init(uint32_t features, uint32_t degree) {
        feature_count = features;
        max_degree = degree;
        exponents.clear();
        if (features == 0) {
            exponents.push_back({});
            return;
        }
        std::vector<uint32_t> cur(features, 0);
        for (uint32_t d = 0; d <= degree; ++d) enumerate_degree(exponents, cur, 0, d);
    }

    std::vector<double> evaluate(const double* f) const {
        std::vector<double> out;
        for (const auto& e : exponents) {
            double v = 1.0;
            for (size_t i = 0; i < e.size(); ++i)
                for (uint32_t k = 0; k < e[i]; ++k) {
                    v *= f[i];
                    if (!std::isfinite(v)) throw std::runtime_error("basis value");
                }
            out.push_back(v);
        }
        return out;
    }
};

struct Scaling {  // FeatureScaling
    double mean = 0, population_variance = 0, scale = 0, zero_scale_threshold = 0;
    bool inactive = false;

    static Scaling fit(const std::vector<double>& v) {
        if (v.empty()) throw std::invalid_argument("empty feature sample");
        NeumaierSum sum;
        double feature_scale = 1.0;
        for (double x : v) {
            if (!std::isfinite(x)) throw std::runtime_error("non-finite feature");
            sum.add(x);
            feature_scale = std::max(feature_scale, std::fabs(x));
        }
        Scaling s;
        s.mean = sum.total() / static_cast<double>(v.size());
        NeumaierSum sq;
        for (double x : v) {
            const double d = x - s.mean;
            sq.add(d * d);
        }
        s.population_variance = sq.total() / static_cast<double>(v.size());
        s.scale = std::sqrt(s.population_variance);
        s.zero_scale_threshold = (64.0 * 2.220446049250313e-16) * feature_scale;
        s.inactive = s.scale <= s.zero_scale_threshold;
        return s;
    }
    double standardize(double x) const { return (x - mean) / scale; }
};

struct Cpqr {  // CpqrFit
    std::vector<size_t> pivot_order;
    std::vector<double> diagonal_abs, coefficients;
    double rank_threshold = 0, residual_sum_squares = 0;
    size_t rank = 0;
};

inline double column_squared_norm(const std::vector<double>& m, size_t rows, size_t cols, size_t start, size_t c) {
    NeumaierSum n;
    for (size_t r = start; r < rows; ++r) {
        const double v = m[r * cols + c];
        n.add(v * v);
    }
    return n.total();
}

inline double column_norm(const std::vector<double>& m, size_t rows, size_t cols, size_t start, size_t c) {
    double scale = 0.0, sum_squares = 1.0;
    for (size_t r = start; r < rows; ++r) {
        const double v = std::fabs(m[r * cols + c]);
        if (v == 0.0) continue;
        if (scale < v) {
            const double ratio = scale / v;
            sum_squares = 1.0 + sum_squares * (ratio * ratio);
            scale = v;
        } else {
            const double ratio = v / scale;
            sum_squares += ratio * ratio;
        }
    }
    return scale == 0.0 ? 0.0 : scale * std::sqrt(sum_squares);
}

inline Cpqr fit_cpqr(const std::vector<double>& matrix, size_t rows, size_t cols, const std::vector<double>& target,
                     double abs_tol, double rel_tol) {
    if (rows == 0 || cols == 0) throw std::invalid_argument("empty QR matrix");
    std::vector<double> qr = matrix, b = target;
    std::vector<size_t> perm(cols);
    for (size_t i = 0; i < cols; ++i) perm[i] = i;
    Cpqr fit;
    const size_t diag_count = std::min(rows, cols);
    for (size_t p = 0; p < diag_count; ++p) {
        size_t sel = p;
        double sel_norm = -1.0;
        for (size_t c = p; c < cols; ++c) {
            const double sq = column_squared_norm(qr, rows, cols, p, c);
            if (sq > sel_norm || (sq == sel_norm && perm[c] < perm[sel])) {
                sel = c;
                sel_norm = sq;
            }
        }
        if (sel != p) {
            for (size_t r = 0; r < rows; ++r) std::swap(qr[r * cols + p], qr[r * cols + sel]);
            std::swap(perm[p], perm[sel]);
        }
        const double norm = column_norm(qr, rows, cols, p, p);
        if (norm == 0.0) {
            qr[p * cols + p] = 0.0;
            fit.diagonal_abs.push_back(0.0);
            continue;
        }
        const size_t di = p * cols + p;
        const double leading = qr[di];
        const double alpha = leading >= 0.0 ? -norm : norm;
        const double denominator = leading - alpha;
        const double tau = (alpha - leading) / alpha;
        if (!std::isfinite(denominator) || !std::isfinite(tau)) throw std::runtime_error("Householder reflector");
        qr[di] = alpha;
        for (size_t r = p + 1; r < rows; ++r) qr[r * cols + p] /= denominator;
        for (size_t c = p + 1; c < cols; ++c) {
            NeumaierSum dot;
            dot.add(qr[p * cols + c]);
            for (size_t r = p + 1; r < rows; ++r) dot.add(qr[r * cols + p] * qr[r * cols + c]);
            const double projection = tau * dot.total();
            qr[p * cols + c] -= projection;
            for (size_t r = p + 1; r < rows; ++r) qr[r * cols + c] -= qr[r * cols + p] * projection;
        }
        NeumaierSum tdot;
        tdot.add(b[p]);
        for (size_t r = p + 1; r < rows; ++r) tdot.add(qr[r * cols + p] * b[r]);
        const double tp = tau * tdot.total();
        b[p] -= tp;
        for (size_t r = p + 1; r < rows; ++r) b[r] -= qr[r * cols + p] * tp;
        fit.diagonal_abs.push_back(std::fabs(alpha));
    }
    const double first = fit.diagonal_abs.empty() ? 0.0 : fit.diagonal_abs[0];
    fit.rank_threshold = std::max(abs_tol, rel_tol * first);
    size_t rank = 0;
    while (rank < fit.diagonal_abs.size() && fit.diagonal_abs[rank] > fit.rank_threshold) ++rank;
    fit.rank = rank;
    std::vector<double> pc(cols, 0.0);
    for (size_t r = rank; r-- > 0;) {
        NeumaierSum up;
        for (size_t c = r + 1; c < rank; ++c) up.add(qr[r * cols + c] * pc[c]);
        const double coefficient = (b[r] - up.total()) / qr[r * cols + r];
        if (!std::isfinite(coefficient)) throw std::runtime_error("back substitution");
        pc[r] = coefficient;
    }
    fit.coefficients.assign(cols, 0.0);
    for (size_t p = 0; p < cols; ++p) fit.coefficients[perm[p]] = pc[p];
    NeumaierSum rss;
    for (size_t r = 0; r < rows; ++r) {
        NeumaierSum pred;
        for (size_t c = 0; c < cols; ++c) pred.add(matrix[r * cols + c] * fit.coefficients[c]);
        const double res = target[r] - pred.total();
        rss.add(res * res);
    }
    fit.residual_sum_squares = rss.total();
    fit.pivot_order = perm;
    return fit;
}

// Least squares through the normal equations with a pivoted Cholesky that
// follows fit_cpqr's pivot and rank rules: at step p the remaining column of
// largest residual squared norm (the Schur complement's diagonal; ties to the
// smaller original index) is chosen, |R_pp| is the square root of that
// diagonal, the rank keeps the leading |R_pp| above max(abs, rel |R_00|) and
// the solution uses the first `rank` pivoted columns only.  In exact
// arithmetic the pivots, |R_pp| and coefficients equal fit_cpqr's; the
// rounding differs (the Gram squares the condition number).  No reference
// counterpart: the device solves LSM regressions this way.
inline Cpqr fit_gram_cholesky(const std::vector<double>& gram, const std::vector<double>& rhs, double yy,
                              size_t rows, size_t cols, double abs_tol, double rel_tol) {
    if (rows == 0 || cols == 0) throw std::invalid_argument("empty regression");
    std::vector<double> s = gram, b = rhs;
    std::vector<size_t> perm(cols);
    for (size_t i = 0; i < cols; ++i) perm[i] = i;
    std::vector<double> l(cols * cols, 0.0);
    Cpqr fit;
    const size_t diag_count = std::min(rows, cols);
    for (size_t p = 0; p < diag_count; ++p) {
        size_t sel = p;
        for (size_t c = p + 1; c < cols; ++c) {
            const double a = s[c * cols + c], m = s[sel * cols + sel];
            if (a > m || (a == m && perm[c] < perm[sel])) sel = c;
        }
        if (sel != p) {
            for (size_t k = 0; k < cols; ++k) std::swap(s[p * cols + k], s[sel * cols + k]);
            for (size_t k = 0; k < cols; ++k) std::swap(s[k * cols + p], s[k * cols + sel]);
            for (size_t k = 0; k < p; ++k) std::swap(l[p * cols + k], l[sel * cols + k]);
            std::swap(b[p], b[sel]);
            std::swap(perm[p], perm[sel]);
        }
        const double d = s[p * cols + p];
        if (!(d > 0.0)) {
            fit.diagonal_abs.push_back(0.0);
            continue;
        }
        const double r = std::sqrt(d);
        l[p * cols + p] = r;
        for (size_t i = p + 1; i < cols; ++i) l[i * cols + p] = s[i * cols + p] / r;
        for (size_t i = p + 1; i < cols; ++i)
            for (size_t j = p + 1; j < cols; ++j) s[i * cols + j] -= l[i * cols + p] * l[j * cols + p];
        fit.diagonal_abs.push_back(r);
    }
    const double first = fit.diagonal_abs.empty() ? 0.0 : fit.diagonal_abs[0];
    fit.rank_threshold = std::max(abs_tol, rel_tol * first);
    size_t rank = 0;
    while (rank < fit.diagonal_abs.size() && fit.diagonal_abs[rank] > fit.rank_threshold) ++rank;
    fit.rank = rank;
    // L_r L_r^T c = b_r: forward then back substitution
    std::vector<double> z(rank, 0.0), c(cols, 0.0);
    for (size_t i = 0; i < rank; ++i) {
        double v = b[i];
        for (size_t k = 0; k < i; ++k) v -= l[i * cols + k] * z[k];
        z[i] = v / l[i * cols + i];
    }
    for (size_t i = rank; i-- > 0;) {
        double v = z[i];
        for (size_t k = i + 1; k < rank; ++k) v -= l[k * cols + i] * c[k];
        c[i] = v / l[i * cols + i];
        if (!std::isfinite(c[i])) throw std::runtime_error("Cholesky back substitution");
    }
    fit.coefficients.assign(cols, 0.0);
    double explained = 0.0;
    for (size_t p = 0; p < rank; ++p) {
        fit.coefficients[perm[p]] = c[p];
        explained += c[p] * b[p];
    }
    fit.residual_sum_squares = yy - explained;
    fit.pivot_order = perm;
    return fit;
}

// Gram matrix, X^T y and y^T y of a row-major design with Neumaier sums.
inline Cpqr fit_normal_equations(const std::vector<double>& design, size_t rows, size_t cols,
                                 const std::vector<double>& target, double abs_tol, double rel_tol) {
    std::vector<double> gram(cols * cols), rhs(cols);
    for (size_t i = 0; i < cols; ++i) {
        for (size_t j = i; j < cols; ++j) {
            NeumaierSum g;
            for (size_t r = 0; r < rows; ++r) g.add(design[r * cols + i] * design[r * cols + j]);
            gram[i * cols + j] = gram[j * cols + i] = g.total();
        }
        NeumaierSum t;
        for (size_t r = 0; r < rows; ++r) t.add(design[r * cols + i] * target[r]);
        rhs[i] = t.total();
    }
    NeumaierSum yy;
    for (size_t r = 0; r < rows; ++r) yy.add(target[r] * target[r]);
    return fit_gram_cholesky(gram, rhs, yy.total(), rows, cols, abs_tol, rel_tol);
}

// Weighted Laguerre functions exp(-x/2) L_k(x), k = 0..degree, by the
// three-term recurrence (k+1) L_{k+1} = (2k+1-x) L_k - k L_{k-1}.  The
// Longstaff--Schwartz basis for the American put; x = spot / strike.
inline void weighted_laguerre(double x, uint32_t degree, double* out) {
    const double w = std::exp(-0.5 * x);
    double prev = 0.0, cur = 1.0;
    for (uint32_t k = 0; k <= degree; ++k) {
        out[k] = w * cur;
        const double next = ((2.0 * k + 1.0 - x) * cur - static_cast<double>(k) * prev) / (k + 1.0);
        prev = cur;
        cur = next;
    }
}

struct Regression {  // PolynomialRegressionModel
    Basis basis;
    std::vector<Scaling> scalings;
    std::vector<size_t> active, pre_excluded, pivot_order;
    std::vector<double> diagonal_abs, coefficients;
    double rank_threshold = 0, residual_sum_squares = 0;
    size_t rank = 0;

    std::vector<double> standardized(const double* f) const {
        std::vector<double> s(scalings.size());
        for (size_t i = 0; i < scalings.size(); ++i) {
            if (!std::isfinite(f[i])) throw std::runtime_error("non-finite feature");
            s[i] = scalings[i].inactive ? 0.0 : scalings[i].standardize(f[i]);
        }
        return s;
    }
    double predict(const double* f) const {
        const std::vector<double> s = standardized(f);
        const std::vector<double> b = basis.evaluate(s.data());
        NeumaierSum p;
        for (size_t i = 0; i < b.size(); ++i) p.add(b[i] * coefficients[i]);
        const double v = p.total();
        if (!std::isfinite(v)) throw std::runtime_error("regression prediction");
        return v;
    }
};

inline Regression fit_polynomial_regression(const Basis& basis, const std::vector<double>& features, size_t rows,
                                            const std::vector<double>& target, double abs_tol, double rel_tol) {
    Regression m;
    m.basis = basis;
    const size_t fc = basis.feature_count;
    std::vector<double> col(rows);
    for (size_t f = 0; f < fc; ++f) {
        for (size_t r = 0; r < rows; ++r) col[r] = features[r * fc + f];
        m.scalings.push_back(Scaling::fit(col));
    }
    for (size_t c = 0; c < basis.exponents.size(); ++c) {
        bool dep = false;
        for (size_t f = 0; f < fc; ++f)
            if (basis.exponents[c][f] != 0 && m.scalings[f].inactive) dep = true;
        (dep ? m.pre_excluded : m.active).push_back(c);
    }
    const size_t ac = m.active.size();
    std::vector<double> design(rows * ac);
    for (size_t r = 0; r < rows; ++r) {
        const std::vector<double> s = m.standardized(&features[r * fc]);
        for (size_t a = 0; a < ac; ++a) {
            double v = 1.0;
            const auto& e = basis.exponents[m.active[a]];
            for (size_t i = 0; i < e.size(); ++i)
                for (uint32_t k = 0; k < e[i]; ++k) v *= s[i];
            design[r * ac + a] = v;
        }
    }
    const Cpqr fit = fit_cpqr(design, rows, ac, target, abs_tol, rel_tol);
    for (size_t p : fit.pivot_order) m.pivot_order.push_back(m.active[p]);
    m.diagonal_abs = fit.diagonal_abs;
    m.rank_threshold = fit.rank_threshold;
    m.rank = fit.rank;
    m.coefficients.assign(basis.exponents.size(), 0.0);
    for (size_t a = 0; a < ac; ++a) m.coefficients[m.active[a]] = fit.coefficients[a];
    m.residual_sum_squares = fit.residual_sum_squares;
    return m;
}

// Weighted-Laguerre continuation model on x = spot / strike (the device's
// plain method; no reference counterpart).
struct LaguerreModel {
    uint32_t degree = 3;
    double strike = 1.0;
    Cpqr fit;

    double predict(double spot) const {
        double phi[16];
        weighted_laguerre(spot / strike, degree, phi);
        NeumaierSum p;
        for (uint32_t k = 0; k <= degree; ++k) p.add(phi[k] * fit.coefficients[k]);
        return p.total();
    }
};

// The reference compares immediate and continuation values on every path, so
// an out-of-the-money path whose extrapolated continuation is negative stops
// with a zero payoff.  The Laguerre policy exercises only in the money
// (immediate > ITM tolerance), as Longstaff and Schwartz do.
struct Decision {  // ExerciseDecisionModel
    bool continue_all = false;
    bool laguerre = false;
    bool itm_only = false;   // Longstaff--Schwartz rule for a monomial policy
    double itm_tol = 0.0;
    Regression model;
    LaguerreModel lag;
    size_t itm_rows = 0;

    double continuation(const double* f) const { return laguerre ? lag.predict(f[0]) : model.predict(f); }
    bool should_exercise(double immediate, const double* f) const {
        if (!std::isfinite(immediate)) throw std::runtime_error("non-finite decision value");
        if (continue_all) return false;
        if ((laguerre || itm_only) && !(immediate > itm_tol)) return false;
        return immediate > continuation(f);
    }
};

struct LsmSettings {
    Basis basis;
    double abs_tol = 1e-14, rel_tol = 1e-12, itm_tol = 0.0;
    uint64_t max_elements = UINT64_MAX;
    bool laguerre = false;   // weighted Laguerre + Gram/Cholesky instead of monomials + CPQR
    bool itm_only = false;   // monomials + CPQR, exercising in-the-money paths only
    double strike = 1.0;
};

inline Decision fit_exercise_decision(const LsmSettings& s, const double* features, size_t rows,
                                      const double* immediate, const std::vector<double>& targets) {
    Decision d;
    const size_t fc = s.basis.feature_count;
    size_t itm = 0;
    for (size_t r = 0; r < rows; ++r)
        if (immediate[r] > s.itm_tol) ++itm;
    d.itm_rows = itm;
    if (itm == 0) {
        d.continue_all = true;
        return d;
    }
    if (itm * fc > s.max_elements) throw std::invalid_argument("regression matrix element limit");
    std::vector<double> f(itm * fc), t(itm);
    size_t k = 0;
    for (size_t r = 0; r < rows; ++r) {
        if (immediate[r] <= s.itm_tol) continue;
        for (size_t i = 0; i < fc; ++i) f[k * fc + i] = features[r * fc + i];
        t[k] = targets[r];
        ++k;
    }
    if (s.laguerre) {
        const uint32_t deg = s.basis.max_degree, cols = deg + 1;
        if (fc != 1 || cols > 16) throw std::invalid_argument("Laguerre basis: one feature, degree < 16");
        std::vector<double> design(itm * cols);
        for (size_t r = 0; r < itm; ++r) weighted_laguerre(f[r] / s.strike, deg, &design[r * cols]);
        d.laguerre = true;
        d.itm_tol = s.itm_tol;
        d.lag.degree = deg;
        d.lag.strike = s.strike;
        d.lag.fit = fit_normal_equations(design, itm, cols, t, s.abs_tol, s.rel_tol);
        return d;
    }
    d.model = fit_polynomial_regression(s.basis, f, itm, t, s.abs_tol, s.rel_tol);
    d.itm_only = s.itm_only;
    d.itm_tol = s.itm_tol;
    return d;
}

struct TrainedPolicy {
    std::vector<Decision> decisions;            // one per non-terminal date
    std::vector<double> realized_cashflows;
    std::vector<size_t> stopping_indices;
};

// immediate: [date][path]; features: [date][path][feature] for non-terminal dates.
inline TrainedPolicy train_exercise_policy(const LsmSettings& s, size_t dates, size_t paths,
                                           const std::vector<double>& immediate, const std::vector<double>& features,
                                           const std::vector<double>& discount) {
    if (dates == 0 || paths == 0) throw std::invalid_argument("empty LSM matrices");
    if (paths > s.max_elements) throw std::invalid_argument("regression matrix element limit");
    for (double v : immediate)
        if (!std::isfinite(v) || v < 0.0) throw std::invalid_argument("immediate values must be finite and non-negative");
    const size_t fc = s.basis.feature_count;
    const size_t expiry = dates - 1;
    TrainedPolicy out;
    out.realized_cashflows.assign(immediate.begin() + static_cast<long>(expiry * paths),
                                  immediate.begin() + static_cast<long>(expiry * paths + paths));
    out.stopping_indices.assign(paths, expiry);
    std::vector<Decision> reverse;
    std::vector<double> targets(paths);
    for (size_t d = expiry; d-- > 0;) {
        for (size_t p = 0; p < paths; ++p) {
            targets[p] = out.realized_cashflows[p] * discount[out.stopping_indices[p]] / discount[d];
            if (!std::isfinite(targets[p])) throw std::runtime_error("non-finite continuation target");
        }
        const double* imm = &immediate[d * paths];
        const double* feat = &features[d * paths * fc];
        Decision dec = fit_exercise_decision(s, feat, paths, imm, targets);
        for (size_t p = 0; p < paths; ++p)
            if (dec.should_exercise(imm[p], feat + p * fc)) {
                out.realized_cashflows[p] = imm[p];
                out.stopping_indices[p] = d;
            }
        reverse.push_back(std::move(dec));
    }
    out.decisions.assign(reverse.rbegin(), reverse.rend());
    return out;
}

struct ValuedPolicy {
    std::vector<double> discounted_cashflows;
    std::vector<size_t> stopping_indices;
    std::vector<uint64_t> exercise_counts;
};

inline ValuedPolicy value_exercise_policy(const std::vector<Decision>& decisions, size_t dates, size_t paths,
                                          size_t fc, const std::vector<double>& immediate,
                                          const std::vector<double>& features, const std::vector<double>& discount) {
    const size_t expiry = dates - 1;
    std::vector<double> realized(immediate.begin() + static_cast<long>(expiry * paths),
                                 immediate.begin() + static_cast<long>(expiry * paths + paths));
    ValuedPolicy v;
    v.stopping_indices.assign(paths, expiry);
    for (size_t d = 0; d < expiry; ++d)
        for (size_t p = 0; p < paths; ++p) {
            if (v.stopping_indices[p] != expiry) continue;
            const double imm = immediate[d * paths + p];
            if (decisions[d].should_exercise(imm, &features[(d * paths + p) * fc])) {
                realized[p] = imm;
                v.stopping_indices[p] = d;
            }
        }
    v.discounted_cashflows.resize(paths);
    v.exercise_counts.assign(dates, 0);
    for (size_t p = 0; p < paths; ++p) {
        v.exercise_counts[v.stopping_indices[p]] += 1;
        v.discounted_cashflows[p] = realized[p] * discount[v.stopping_indices[p]];
    }
    return v;
}

}  // namespace oracle
