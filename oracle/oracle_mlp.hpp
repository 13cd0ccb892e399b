// ORACLE -- test infrastructure only.  FP64 restatement of the continuation
// network of the learned LSM policy (no reference counterpart: the
// reference's early-exercise roadmap leaves neural regressions out of scope,
// design/roadmaps/early-exercise-roadmap-v0.1.md:61).
//
// Network: 2 -> 64 -> 64 -> 1, SiLU activations, mean-squared-error loss
//   L = (1/n) sum_r (out_r - y_r)^2.
// Parameter vector (float32 on the device, FP64 here):
//   W1[64][2] | b1[64] | W2[64][64] (row = output unit) | b2[64] | w3[64] | b3
// With `bf16` set, the quantities the device feeds to its BF16 matrix cores
// are rounded to bfloat16 (round to nearest even): W2, both layers' hidden
// activations, both layers' pre-activation gradients, the inputs and the
// output gradient where they enter a contraction over rows; everything else
// stays FP64.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {

constexpr int kMlpH = 64;
constexpr int kMlpW1 = 0, kMlpB1 = 128, kMlpW2 = 192, kMlpB2 = 4288, kMlpW3 = 4352, kMlpB3 = 4416;
constexpr int kMlpParams = 4417;

inline double bf16_round(double v) {
    float f = static_cast<float>(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;   // inf / nan unchanged
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

inline double sigmoid(double a) { return 1.0 / (1.0 + std::exp(-a)); }

// Forward pass of one row; fills the activations when asked.
inline double mlp_forward(const double* p, double x1, double x2, bool bf16, double* a1 = nullptr,
                          double* h1 = nullptr, double* z2 = nullptr, double* h2 = nullptr) {
    double la1[kMlpH], lh1[kMlpH], lz2[kMlpH], lh2[kMlpH];
    for (int f = 0; f < kMlpH; ++f) {
        la1[f] = p[kMlpW1 + 2 * f] * x1 + p[kMlpW1 + 2 * f + 1] * x2 + p[kMlpB1 + f];
        lh1[f] = la1[f] * sigmoid(la1[f]);
    }
    double out = p[kMlpB3];
    for (int o = 0; o < kMlpH; ++o) {
        double z = 0.0;
        for (int f = 0; f < kMlpH; ++f) {
            const double w = bf16 ? bf16_round(p[kMlpW2 + kMlpH * o + f]) : p[kMlpW2 + kMlpH * o + f];
            const double h = bf16 ? bf16_round(lh1[f]) : lh1[f];
            z += w * h;
        }
        lz2[o] = z + p[kMlpB2 + o];
        lh2[o] = lz2[o] * sigmoid(lz2[o]);
        out += p[kMlpW3 + o] * lh2[o];
    }
    if (a1) std::memcpy(a1, la1, sizeof la1);
    if (h1) std::memcpy(h1, lh1, sizeof lh1);
    if (z2) std::memcpy(z2, lz2, sizeof lz2);
    if (h2) std::memcpy(h2, lh2, sizeof lh2);
    return out;
}

// Loss and its gradient over n rows.
inline double mlp_loss_gradient(const double* p, const double* x, const double* y, size_t n, bool bf16,
                                double* grad) {
    std::vector<double> g(kMlpParams, 0.0);
    double loss = 0.0;
    const double inv = 1.0 / static_cast<double>(n);
    for (size_t r = 0; r < n; ++r) {
        double a1[kMlpH], h1[kMlpH], z2[kMlpH], h2[kMlpH];
        const double x1 = x[2 * r], x2 = x[2 * r + 1];
        const double out = mlp_forward(p, x1, x2, bf16, a1, h1, z2, h2);
        const double e = out - y[r];
        loss += e * e * inv;
        const double go = 2.0 * e * inv;
        g[kMlpB3] += go;
        const double gob = bf16 ? bf16_round(go) : go;
        double gz[kMlpH];
        for (int o = 0; o < kMlpH; ++o) {
            g[kMlpW3 + o] += gob * (bf16 ? bf16_round(h2[o]) : h2[o]);
            const double s = sigmoid(z2[o]);
            gz[o] = go * p[kMlpW3 + o] * s * (1.0 + z2[o] * (1.0 - s));
            if (bf16) gz[o] = bf16_round(gz[o]);
            g[kMlpB2 + o] += gz[o];
        }
        for (int o = 0; o < kMlpH; ++o)
            for (int f = 0; f < kMlpH; ++f) g[kMlpW2 + kMlpH * o + f] += gz[o] * (bf16 ? bf16_round(h1[f]) : h1[f]);
        for (int f = 0; f < kMlpH; ++f) {
            double gh = 0.0;
            for (int o = 0; o < kMlpH; ++o)
                gh += (bf16 ? bf16_round(p[kMlpW2 + kMlpH * o + f]) : p[kMlpW2 + kMlpH * o + f]) * gz[o];
            const double s = sigmoid(a1[f]);
            double ga = gh * s * (1.0 + a1[f] * (1.0 - s));
            if (bf16) ga = bf16_round(ga);
            g[kMlpB1 + f] += ga;
            g[kMlpW1 + 2 * f] += ga * (bf16 ? bf16_round(x1) : x1);
            g[kMlpW1 + 2 * f + 1] += ga * (bf16 ? bf16_round(x2) : x2);
        }
    }
    std::memcpy(grad, g.data(), sizeof(double) * kMlpParams);
    return loss;
}

}  // namespace oracle
