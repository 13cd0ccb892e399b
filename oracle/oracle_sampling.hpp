// ORACLE -- test infrastructure only.  CPU FP64 restatement of the reference's
// sampling layer, used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline as the checker.  Nothing in the product links or calls it.
//
// Restates crates/pricing/src/engine/sampling/{random,qmc,bridge}.rs in the
// reference's operation order (no FMA contraction: build with
// -ffp-contract=off).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------- Philox --
// random.rs:4-7 Random123 constants; :100-112 ten rounds, key bumped between
// rounds; :131-143 round function.
constexpr uint32_t kPhiloxM0 = 0xd2511f53u;
constexpr uint32_t kPhiloxM1 = 0xcd9e8d57u;
constexpr uint32_t kPhiloxW0 = 0x9e3779b9u;
constexpr uint32_t kPhiloxW1 = 0xbb67ae85u;

enum Domain : uint32_t {  // random.rs:13-20
    kValuation = 0,
    kLsmTrain = 1,
    kRqmcScramble = 2,
    kDiagnostics = 3,
};

inline void philox4x32_10(uint64_t seed, const uint32_t in[4], uint32_t out[4]) {
    uint32_t key0 = static_cast<uint32_t>(seed);
    uint32_t key1 = static_cast<uint32_t>(seed >> 32);
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = static_cast<uint64_t>(kPhiloxM0) * c0;
        const uint64_t p1 = static_cast<uint64_t>(kPhiloxM1) * c2;
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32);
        const uint32_t n0 = hi1 ^ c1 ^ key0;
        const uint32_t n1 = static_cast<uint32_t>(p1);
        const uint32_t n2 = hi0 ^ c3 ^ key1;
        const uint32_t n3 = static_cast<uint32_t>(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        if (round + 1 < 10) {
            key0 += kPhiloxW0;
            key1 += kPhiloxW1;
        }
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// RandomCoordinate::counter/lane (random.rs:62-76): counter = [path lo, path hi,
// dim/4, domain], lane = dim % 4.
inline uint32_t philox_word(uint64_t seed, uint64_t path, uint32_t dimension, uint32_t domain) {
    const uint32_t counter[4] = {static_cast<uint32_t>(path), static_cast<uint32_t>(path >> 32),
                                 dimension / 4, domain};
    uint32_t block[4];
    philox4x32_10(seed, counter, block);
    return block[dimension % 4];
}

// open_unit_interval (random.rs:147-149).
inline double open_unit_interval(uint32_t word) {
    return (static_cast<double>(word) + 0.5) * (1.0 / 4294967296.0);
}

// Wichura AS241 PPND16 (random.rs:176-277), Horner from the top coefficient.
inline double as241_poly(double x, const double c[8]) {
    double v = c[7];
    for (int i = 6; i >= 0; --i) v = v * x + c[i];
    return v;
}

inline double inverse_standard_normal(double p) {
    static const double A[8] = {3.3871328727963665, 133.14166789178438, 1971.5909503065513, 13731.69376550946, 45921.95393154987, 67265.7709270087, 33430.57558358836, 2509.0809287301227};
    static const double B[8] = {1.0, 42.31333070160091, 687.1870074920579, 5394.196021424751, 21213.794301586597, 39307.89580009271, 28729.085735721943, 5226.495278852855};
    static const double C[8] = {1.4234371107496835, 4.630337846156546, 5.769497221460691, 3.6478483247632045, 1.2704582524523684, 0.2417807251774506, 0.022723844989269185, 0.0007745450142783414};
    static const double D[8] = {1.0, 2.053191626637759, 1.6763848301838038, 0.6897673349851, 0.14810397642748008, 0.015198666563616457, 0.0005475938084995345, 1.0507500716444168e-9};
    static const double E[8] = {6.657904643501104, 5.463784911164114, 1.7848265399172913, 0.2965605718285049, 0.026532189526576123, 0.0012426609473880784, 0.000027115555687434876, 0.00000020103343992922882};
    static const double F[8] = {1.0, 0.5998322065558879, 0.1369298809227358, 0.014875361290850615, 0.0007868691311456133, 0.00001846318317010547, 0.00000014215117583168448, 2.0442631033899397e-15};
    if (!std::isfinite(p) || p <= 0.0 || p >= 1.0)
        throw std::domain_error("normal quantile probability outside (0,1)");
    const double centered = p - 0.5;
    if (std::fabs(centered) <= 0.425) {
        const double a = 0.180625 - centered * centered;
        return centered * as241_poly(a, A) / as241_poly(a, B);
    }
    const double tail = centered < 0.0 ? p : 1.0 - p;
    const double r = std::sqrt(-std::log(tail));
    double magnitude;
    if (r <= 5.0) {
        const double a = r - 1.6;
        magnitude = as241_poly(a, C) / as241_poly(a, D);
    } else {
        const double a = r - 5.0;
        magnitude = as241_poly(a, E) / as241_poly(a, F);
    }
    return centered < 0.0 ? -magnitude : magnitude;
}

inline double philox_normal(uint64_t seed, uint64_t path, uint32_t dimension, uint32_t domain) {
    return inverse_standard_normal(open_unit_interval(philox_word(seed, path, dimension, domain)));
}

// ----------------------------------------------------------------- Sobol --
// Joe--Kuo 6.21201 32-bit directions, dimension-major (qmc.rs:41-83).  The
// table is the generated direction asset (SHA-256 pinned by the tests).
constexpr uint32_t kMaxSobolDimension = 21201;

struct DirectionTable {
    std::vector<uint32_t> words;  // [dimension][bit]
    uint32_t dimensions = 0;

    const uint32_t* dimension(uint32_t d) const { return &words[static_cast<size_t>(d) * 32]; }
};

inline uint32_t sobol_word(const uint32_t* directions, uint32_t point) {
    uint32_t gray = point ^ (point >> 1);
    uint32_t word = 0;
    int bit = 0;
    while (gray != 0) {
        if (gray & 1u) word ^= directions[bit];
        gray >>= 1;
        ++bit;
    }
    return word;
}

struct Scramble32 {  // qmc.rs:107-133
    uint32_t shift;
    uint32_t rows[32];

    uint32_t apply(uint32_t input) const {
        uint32_t output = 0;
        for (int row = 0; row < 32; ++row) {
            const uint32_t parity = static_cast<uint32_t>(__builtin_popcount(input & rows[row])) & 1u;
            output |= parity << (31 - row);
        }
        return output ^ shift;
    }
};

// compile_scramble (qmc.rs:346-369): Philox domain RqmcScramble, path =
// scramble index, dimension = 32*dim + row.
inline Scramble32 compile_scramble(uint64_t seed, uint64_t scramble, uint32_t dimension) {
    Scramble32 s{};
    const uint32_t block = dimension * 32u;
    s.shift = philox_word(seed, scramble, block, kRqmcScramble);
    s.rows[0] = 0x80000000u;
    for (uint32_t row = 1; row < 32; ++row) {
        const uint32_t raw = philox_word(seed, scramble, block + row, kRqmcScramble);
        const uint32_t before = 0xffffffffu << (32 - row);
        const uint32_t diagonal = 1u << (31 - row);
        s.rows[row] = (raw & before) | diagonal;
    }
    return s;
}

struct RqmcPlan {  // qmc.rs:190-344
    uint64_t points = 0;
    uint32_t scrambles = 0;
    uint32_t dimension = 0;
    const DirectionTable* table = nullptr;
    std::vector<Scramble32> scramble_table;  // [scramble][dimension]

    void compile(const DirectionTable& t, uint64_t points_per_scramble, uint32_t scramble_count,
                 uint64_t seed, uint32_t effective_dimension) {
        if (effective_dimension == 0) throw std::invalid_argument("zero effective Sobol dimension");
        if (effective_dimension > kMaxSobolDimension || effective_dimension > t.dimensions)
            throw std::invalid_argument("Sobol dimension limit exceeded");
        table = &t;
        points = points_per_scramble;
        scrambles = scramble_count;
        dimension = effective_dimension;
        scramble_table.resize(static_cast<size_t>(scramble_count) * effective_dimension);
        for (uint32_t s = 0; s < scramble_count; ++s)
            for (uint32_t d = 0; d < effective_dimension; ++d)
                scramble_table[static_cast<size_t>(s) * effective_dimension + d] = compile_scramble(seed, s, d);
    }

    uint32_t word(uint32_t scramble, uint64_t point, uint32_t d) const {
        const uint32_t raw = sobol_word(table->dimension(d), static_cast<uint32_t>(point));
        return scramble_table[static_cast<size_t>(scramble) * dimension + d].apply(raw);
    }

    double uniform(uint32_t scramble, uint64_t point, uint32_t d) const {
        return open_unit_interval(word(scramble, point, d));
    }
};

// ------------------------------------------------------- Brownian bridge --
// bridge.rs:164-363 (plan, apply, reverse) and :395-467 (instruction order:
// largest conditional variance, ties to the earlier time, then lower node).
struct BridgeInstruction {
    uint32_t node, left, right, rank;
    double left_weight, right_weight, standard_deviation;
    bool terminal;
};

struct BridgePlan {
    std::vector<double> times;
    std::vector<double> inverse_sqrt_dt;
    std::vector<BridgeInstruction> instructions;

    struct Candidate {
        double variance, time;
        size_t node, left, right;
    };
    // Ord: variance (total order), then earlier time is greater, then lower node is greater.
    static bool less(const Candidate& a, const Candidate& b) {
        if (a.variance != b.variance) return a.variance < b.variance;
        if (a.time != b.time) return a.time > b.time;
        return a.node > b.node;
    }

    static bool interval_candidate(const std::vector<double>& t, size_t left, size_t right,
                                   Candidate* out) {
        if (right <= left + 1) return false;
        const double denominator = t[right] - t[left];
        bool have = false;
        Candidate best{};
        for (size_t node = left + 1; node < right; ++node) {
            const double variance = ((t[node] - t[left]) * (t[right] - t[node])) / denominator;
            if (!std::isfinite(variance) || variance <= 0.0)
                throw std::invalid_argument("invalid Brownian bridge conditional variance");
            Candidate c{variance, t[node], node, left, right};
            // Iterator::max keeps the last maximum among equals (Ord::max returns `other` on ties).
            if (!have || !less(c, best)) { best = c; have = true; }
        }
        *out = best;
        return true;
    }

    void compile(const std::vector<double>& t) {
        if (t.size() < 2) throw std::invalid_argument("Brownian bridge needs at least two nodes");
        for (double x : t)
            if (!std::isfinite(x)) throw std::invalid_argument("non-finite bridge time");
        if (t[0] != 0.0 || std::signbit(t[0])) throw std::invalid_argument("bridge must start at +0");
        for (size_t i = 0; i + 1 < t.size(); ++i)
            if (!(t[i + 1] > t[i])) throw std::invalid_argument("bridge times must increase");
        times = t;
        inverse_sqrt_dt.clear();
        for (size_t i = 0; i + 1 < t.size(); ++i) inverse_sqrt_dt.push_back(1.0 / std::sqrt(t[i + 1] - t[i]));
        instructions.clear();
        const size_t terminal = t.size() - 1;
        instructions.push_back({static_cast<uint32_t>(terminal), 0, 0, 0, 0.0, 0.0, std::sqrt(t[terminal]), true});
        auto cmp = [](const Candidate& a, const Candidate& b) { return less(a, b); };
        std::priority_queue<Candidate, std::vector<Candidate>, decltype(cmp)> heap(cmp);
        Candidate c;
        if (interval_candidate(t, 0, terminal, &c)) heap.push(c);
        uint32_t rank = 1;
        while (!heap.empty()) {
            c = heap.top();
            heap.pop();
            const double denominator = t[c.right] - t[c.left];
            const double lw = (t[c.right] - t[c.node]) / denominator;
            const double rw = (t[c.node] - t[c.left]) / denominator;
            instructions.push_back({static_cast<uint32_t>(c.node), static_cast<uint32_t>(c.left),
                                    static_cast<uint32_t>(c.right), rank, lw, rw, std::sqrt(c.variance), false});
            ++rank;
            Candidate l, r;
            if (interval_candidate(t, c.left, c.node, &l)) heap.push(l);
            if (interval_candidate(t, c.node, c.right, &r)) heap.push(r);
        }
    }

    std::vector<double> apply_one_factor(const std::vector<double>& normals) const {
        if (normals.size() != inverse_sqrt_dt.size()) throw std::invalid_argument("bridge normal count");
        std::vector<double> w(times.size(), 0.0);
        for (const auto& in : instructions) {
            if (in.terminal) {
                w[in.node] = in.standard_deviation * normals[in.rank];
            } else {
                w[in.node] = in.left_weight * w[in.left] + in.right_weight * w[in.right] +
                             in.standard_deviation * normals[in.rank];
            }
        }
        std::vector<double> out(inverse_sqrt_dt.size());
        for (size_t i = 0; i < out.size(); ++i) out[i] = (w[i + 1] - w[i]) * inverse_sqrt_dt[i];
        return out;
    }

    std::vector<double> reverse_one_factor(const std::vector<double>& increment_adjoints) const {
        if (increment_adjoints.size() != inverse_sqrt_dt.size())
            throw std::invalid_argument("bridge increment count");
        std::vector<double> va(times.size(), 0.0);
        for (size_t i = 0; i < increment_adjoints.size(); ++i) {
            const double scaled = increment_adjoints[i] * inverse_sqrt_dt[i];
            va[i] -= scaled;
            va[i + 1] += scaled;
        }
        std::vector<double> na(inverse_sqrt_dt.size(), 0.0);
        for (auto it = instructions.rbegin(); it != instructions.rend(); ++it) {
            if (it->terminal) {
                na[it->rank] += va[it->node] * it->standard_deviation;
            } else {
                const double adjoint = va[it->node];
                na[it->rank] += adjoint * it->standard_deviation;
                va[it->left] += adjoint * it->left_weight;
                va[it->right] += adjoint * it->right_weight;
            }
        }
        return na;
    }
};

}  // namespace oracle
