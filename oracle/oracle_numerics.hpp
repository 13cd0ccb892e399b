// ORACLE -- test infrastructure only.  Restates the reference's deterministic
// reductions and fixed-block executor:
//   crates/pricing-numerics/src/reduction.rs:3-161, 244-271
//   crates/pricing/src/engine/mc/execution.rs:146-166, 361-744
#pragma once

#include <atomic>
#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <thread>
#include <vector>

namespace oracle {

struct NeumaierSum {  // reduction.rs:3-51
    double sum = 0.0;
    double correction = 0.0;

    void add(double value) {
        const double updated = sum + value;
        if (std::fabs(sum) >= std::fabs(value)) {
            correction += (sum - updated) + value;
        } else {
            correction += (value - updated) + sum;
        }
        sum = updated;
    }
    double total() const { return sum + correction; }
    static NeumaierSum merged(const NeumaierSum& l, const NeumaierSum& r) {
        NeumaierSum m;
        m.add(l.sum);
        m.add(l.correction);
        m.add(r.sum);
        m.add(r.correction);
        return m;
    }
};

struct CenteredMoment {  // reduction.rs:71-161 (Welford add, Chan-Golub-LeVeque merge)
    uint64_t count = 0;
    double mean = 0.0;
    double m2 = 0.0;

    void add(double value) {
        const uint64_t next = count + 1;
        const double delta = value - mean;
        const double next_mean = mean + delta / static_cast<double>(next);
        m2 += delta * (value - next_mean);
        mean = next_mean;
        count = next;
    }
    static CenteredMoment merged(const CenteredMoment& l, const CenteredMoment& r) {
        if (l.count == 0) return r;
        if (r.count == 0) return l;
        CenteredMoment m;
        m.count = l.count + r.count;
        const double delta = r.mean - l.mean;
        const double lw = static_cast<double>(l.count);
        const double rw = static_cast<double>(r.count);
        const double n = static_cast<double>(m.count);
        m.mean = l.mean + delta * (rw / n);
        const double between = delta * delta * (lw * rw / n);
        m.m2 = (l.m2 + r.m2) + between;
        return m;
    }
    static CenteredMoment two_pass(const std::vector<double>& values) {
        CenteredMoment m;
        if (values.empty()) return m;
        NeumaierSum s;
        for (double v : values) s.add(v);
        m.mean = s.total() / static_cast<double>(values.size());
        NeumaierSum q;
        for (double v : values) {
            const double d = v - m.mean;
            q.add(d * d);
        }
        m.m2 = q.total();
        m.count = values.size();
        return m;
    }
    bool has_variance() const { return count > 1; }
    double sample_variance() const { return m2 / static_cast<double>(count - 1); }
};

template <typename T, typename Merge>
T fixed_tree_reduce(std::vector<T> level, Merge merge) {  // reduction.rs:259-271
    if (level.empty()) return T{};
    while (level.size() > 1) {
        std::vector<T> next;
        next.reserve((level.size() + 1) / 2);
        size_t i = 0;
        for (; i + 1 < level.size(); i += 2) next.push_back(merge(level[i], level[i + 1]));
        if (i < level.size()) next.push_back(level[i]);
        level.swap(next);
    }
    return level[0];
}

struct Statistics {  // DeterministicStatistics (execution.rs:136-166)
    NeumaierSum sum;
    CenteredMoment moments;

    static Statistics two_pass(const std::vector<double>& values) {
        Statistics s;
        for (double v : values) s.sum.add(v);
        s.moments = CenteredMoment::two_pass(values);
        return s;
    }
};

// Per-unit evaluation failure (TryExecutionError::Evaluation, execution.rs:168-196).
struct EvaluationError : std::runtime_error {
    uint64_t sampling_unit;
    EvaluationError(uint64_t unit, const std::string& what)
        : std::runtime_error(what), sampling_unit(unit) {}
};

// try_map_reduce_statistics_array_* (execution.rs:550-727): fixed logical
// blocks, each reduced sequentially (Neumaier + Welford) in unit order, then a
// balanced tree over blocks.  Worker count changes scheduling only.
template <int N>
std::vector<Statistics> map_reduce_statistics(
    uint64_t units, uint64_t block_size, unsigned workers,
    const std::function<void(uint64_t, double*)>& evaluate);

// The same executor for a run-time number of components per unit.
inline std::vector<Statistics> map_reduce_statistics_n(
    uint64_t units, uint64_t block_size, unsigned workers, int width,
    const std::function<void(uint64_t, double*)>& evaluate) {
    if (block_size == 0) throw std::invalid_argument("zero reduction block size");
    const int N = width;
    const uint64_t block_count = (units + block_size - 1) / block_size;
    std::vector<std::vector<Statistics>> partial(block_count, std::vector<Statistics>(N));
    std::atomic<uint64_t> next{0};
    std::atomic<bool> failed{false};
    uint64_t failing_unit = UINT64_MAX;
    std::string failure;
    std::mutex* guard = new std::mutex();
    auto work = [&]() {
        std::vector<double> values(N);
        for (;;) {
            const uint64_t b = next.fetch_add(1);
            if (b >= block_count || failed.load()) return;
            const uint64_t begin = b * block_size;
            const uint64_t end = std::min(units, begin + block_size);
            auto& stats = partial[b];
            for (uint64_t u = begin; u < end; ++u) {
                for (int c = 0; c < N; ++c) values[c] = 0.0;
                try {
                    evaluate(u, values.data());
                } catch (const std::exception& e) {
                    std::lock_guard<std::mutex> lock(*guard);
                    if (u < failing_unit) { failing_unit = u; failure = e.what(); }
                    failed.store(true);
                    return;
                }
                for (int c = 0; c < N; ++c) {
                    stats[c].sum.add(values[c]);
                    stats[c].moments.add(values[c]);
                }
            }
        }
    };
    const unsigned n = std::max(1u, workers);
    if (n == 1) {
        work();
    } else {
        std::vector<std::thread> pool;
        for (unsigned i = 0; i < n; ++i) pool.emplace_back(work);
        for (auto& t : pool) t.join();
    }
    delete guard;
    if (failed.load()) throw EvaluationError(failing_unit, failure);
    std::vector<Statistics> out(N);
    for (int c = 0; c < N; ++c) {
        std::vector<NeumaierSum> sums;
        std::vector<CenteredMoment> moments;
        sums.reserve(block_count);
        moments.reserve(block_count);
        for (uint64_t b = 0; b < block_count; ++b) {
            sums.push_back(partial[b][c].sum);
            moments.push_back(partial[b][c].moments);
        }
        out[c].sum = fixed_tree_reduce(sums, NeumaierSum::merged);
        out[c].moments = fixed_tree_reduce(moments, CenteredMoment::merged);
    }
    return out;
}

template <int N>
std::vector<Statistics> map_reduce_statistics(
    uint64_t units, uint64_t block_size, unsigned workers,
    const std::function<void(uint64_t, double*)>& evaluate) {
    return map_reduce_statistics_n(units, block_size, workers, N, evaluate);
}

}  // namespace oracle
