// ORACLE -- test infrastructure only.  Restates the reference's lifted
// Heston rough-volatility path scheme in FP64:
//   Lanczos gamma on [1/2, 2]     crates/pricing-numerics/src/fractional.rs:1-30
//   LiftedHeston::new/from_rough  crates/pricing/src/models/rough_volatility.rs:180-212, 564-606
//   random layout, lifted variance, asset path
//                                 crates/pricing/src/engine/processes/rough_volatility/mod.rs:167-175,
//                                 206-238, 354-392, 505-547
#pragma once

#include <cmath>
#include <stdexcept>
#include <vector>

#include "oracle_numerics.hpp"

namespace oracle {

// gamma_half_to_two: Lanczos (g = 7, nine terms; Godfrey's published
// coefficients) on [1/2, 2].
inline double gamma_half_to_two(double x) {
    if (!std::isfinite(x) || !(x >= 0.5 && x <= 2.0)) throw std::invalid_argument("gamma argument outside [1/2, 2]");
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

struct LiftedHeston {
    double v0 = 0, kappa = 0, theta = 0, nu = 0, rho = 0;
    std::vector<double> weights, rates;

    void validate() const {
        for (double v : {v0, kappa, theta, nu})
            if (!std::isfinite(v) || v < 0.0) throw std::invalid_argument("lifted Heston parameter must be non-negative");
        if (!std::isfinite(rho) || rho < -1.0 || rho > 1.0) throw std::invalid_argument("lifted Heston correlation");
        if (weights.empty() || weights.size() != rates.size() || weights.size() > 256)
            throw std::invalid_argument("lifted Heston factor count");
        for (size_t i = 0; i < weights.size(); ++i) {
            if (!std::isfinite(weights[i]) || weights[i] < 0.0 || !std::isfinite(rates[i]) || rates[i] < 0.0)
                throw std::invalid_argument("lifted Heston weight/rate");
            if (weights[i] == 0.0) throw std::invalid_argument("lifted Heston zero weight");
        }
    }

    // LiftedHeston::from_rough: geometric-bin quadrature of the fractional
    // kernel's Laplace measure.
    void from_rough(double hurst, size_t factors, double ratio) {
        if (factors == 0 || factors > 256 || !std::isfinite(ratio) || ratio <= 1.0)
            throw std::invalid_argument("lifted Heston quadrature");
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
        const double denominator = gamma_half_to_two(alpha) * gamma_half_to_two(2.0 - alpha);
        const double mass_factor = std::expm1(p * log_ratio) / denominator;
        const double rate_factor = p / (1.0 + p) * std::expm1((1.0 + p) * log_ratio) / std::expm1(p * log_ratio);
        for (size_t j = 0; j < factors; ++j) {
            const double log_lower = (static_cast<double>(j) - static_cast<double>(factors) / 2.0) * log_ratio;
            weights.push_back(mass_factor * std::exp(p * log_lower));
            rates.push_back(rate_factor * std::exp(log_lower));
        }
    }
};

// lifted_variance + asset_path (beta = 1).  normals: 2n values, spot block
// then independent variance block.  Returns the forwards at every node.
inline void lifted_heston_path(const LiftedHeston& m, const std::vector<double>& times, double initial,
                               const double* z, std::vector<double>& forwards, std::vector<double>& variance,
                               std::vector<double>& factors) {
    const size_t n = times.size() - 1;
    factors.assign(m.weights.size(), 0.0);
    variance.assign(1, m.v0);
    const double independent = std::sqrt(std::max(1.0 - m.rho * m.rho, 0.0));
    for (size_t j = 0; j < n; ++j) {
        const double dt = times[j + 1] - times[j];
        const double dw = std::sqrt(dt) * (m.rho * z[j] + independent * z[n + j]);
        const double common = m.kappa * (m.theta - variance[j]) * dt + m.nu * std::sqrt(variance[j]) * dw;
        NeumaierSum sum;
        sum.add(m.v0);
        for (size_t i = 0; i < factors.size(); ++i) {
            const double denominator = 1.0 + m.rates[i] * dt;
            if (!std::isfinite(denominator)) throw std::runtime_error("lifted time scale overflow");
            const double f = (factors[i] + common) / denominator;
            if (!std::isfinite(f)) throw std::runtime_error("rough path overflow");
            factors[i] = f;
            sum.add(m.weights[i] * f);
        }
        const double value = sum.total();
        if (!std::isfinite(value)) throw std::runtime_error("rough path overflow");
        variance.push_back(std::max(value, 0.0));
    }
    forwards.assign(1, initial);
    for (size_t j = 0; j < n; ++j) {
        const double dt = times[j + 1] - times[j];
        const double prev = forwards[j];
        const double v = variance[j];
        if (!std::isfinite(v) || v < 0.0) throw std::runtime_error("rough variance");
        const double noise = std::sqrt(v) * std::sqrt(dt) * z[j];
        const double next = prev * std::exp(-0.5 * v * dt + noise);
        if (next <= 0.0) throw std::runtime_error("rough asset underflow");
        if (!std::isfinite(next)) throw std::runtime_error("rough path overflow");
        forwards.push_back(next);
    }
}

}  // namespace oracle
