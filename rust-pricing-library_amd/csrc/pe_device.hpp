// Device primitives for the gfx950 path kernels: counter-based Philox,
// the AS241 inverse normal in FP64 and FP32, and compensated accumulators.
//
// Philox4x32-10 with Random123 constants and the reference's coordinate ABI
// (crates/pricing/src/engine/sampling/random.rs:62-143); AS241 PPND16 with the
// reference's coefficient table (random.rs:176-277).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pe {

#define PE_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------- Philox --
struct U4 {
    uint32_t x, y, z, w;
};

PE_HD U4 philox4x32_10(uint64_t seed, U4 c) {
    uint32_t k0 = static_cast<uint32_t>(seed);
    uint32_t k1 = static_cast<uint32_t>(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(0xd2511f53u) * c.x;
        const uint64_t p1 = static_cast<uint64_t>(0xcd9e8d57u) * c.z;
        const U4 n{static_cast<uint32_t>(p1 >> 32) ^ c.y ^ k0, static_cast<uint32_t>(p1),
                   static_cast<uint32_t>(p0 >> 32) ^ c.w ^ k1, static_cast<uint32_t>(p0)};
        c = n;
        k0 += 0x9e3779b9u;  // the bump after the last round is never used
        k1 += 0xbb67ae85u;
    }
    return c;
}

// Words for dimensions 4q..4q+3 of one sampling unit: one Philox block serves
// four consecutive dimensions (lane = dim % 4).
PE_HD U4 philox_block(uint64_t seed, uint64_t path, uint32_t quad, uint32_t domain) {
    return philox4x32_10(seed, U4{static_cast<uint32_t>(path), static_cast<uint32_t>(path >> 32), quad, domain});
}

// The same block for a wave-uniform seed, with the ten rounds' keys formed
// here on the scalar unit: the seed passes through an empty asm statement,
// so the key schedule is not hoisted out of the caller's loops into twenty
// scalar registers that spill into vector lanes.  Bit-identical words.
__device__ __forceinline__ U4 philox_block_uniform(uint64_t seed, uint64_t path, uint32_t quad, uint32_t domain) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(seed));
#endif
    return philox_block(seed, path, quad, domain);
}

PE_HD uint32_t lane_of(const U4& b, uint32_t lane) {
    return lane == 0 ? b.x : lane == 1 ? b.y : lane == 2 ? b.z : b.w;
}

// ------------------------------------------------------------------ AS241 --
// Reference coefficient table (random.rs:181-240), kept digit for digit, as
// literals so both the FP64 and FP32 evaluations fold them into immediates.
#define PE_H8(x, c0, c1, c2, c3, c4, c5, c6, c7) \
    ((((((((c7) * (x) + (c6)) * (x) + (c5)) * (x) + (c4)) * (x) + (c3)) * (x) + (c2)) * (x) + (c1)) * (x) + (c0))
#define PE_AS241_A(T, x) PE_H8(x, T(3.3871328727963665), T(133.14166789178438), T(1971.5909503065513), \
    T(13731.69376550946), T(45921.95393154987), T(67265.7709270087), T(33430.57558358836), T(2509.0809287301227))
#define PE_AS241_B(T, x) PE_H8(x, T(1.0), T(42.31333070160091), T(687.1870074920579), T(5394.196021424751), \
    T(21213.794301586597), T(39307.89580009271), T(28729.085735721943), T(5226.495278852855))
#define PE_AS241_C(T, x) PE_H8(x, T(1.4234371107496835), T(4.630337846156546), T(5.769497221460691), \
    T(3.6478483247632045), T(1.2704582524523684), T(0.2417807251774506), T(0.022723844989269185), \
    T(0.0007745450142783414))
#define PE_AS241_D(T, x) PE_H8(x, T(1.0), T(2.053191626637759), T(1.6763848301838038), T(0.6897673349851), \
    T(0.14810397642748008), T(0.015198666563616457), T(0.0005475938084995345), T(1.0507500716444168e-9))
#define PE_AS241_E(T, x) PE_H8(x, T(6.657904643501104), T(5.463784911164114), T(1.7848265399172913), \
    T(0.2965605718285049), T(0.026532189526576123), T(0.0012426609473880784), T(0.000027115555687434876), \
    T(0.00000020103343992922882))
#define PE_AS241_F(T, x) PE_H8(x, T(1.0), T(0.5998322065558879), T(0.1369298809227358), \
    T(0.014875361290850615), T(0.0007868691311456133), T(0.00001846318317010547), \
    T(0.00000014215117583168448), T(2.0442631033899397e-15))

// FP64: exactly the reference's arithmetic on the midpoint probability.
PE_HD double normal_from_word_f64(uint32_t word) {
    const double p = (static_cast<double>(word) + 0.5) * (1.0 / 4294967296.0);
    const double centered = p - 0.5;
    if (fabs(centered) <= 0.425) {
        const double a = 0.180625 - centered * centered;
        return centered * PE_AS241_A(double, a) / PE_AS241_B(double, a);
    }
    const double tail = centered < 0.0 ? p : 1.0 - p;
    const double r = sqrt(-log(tail));
    double m;
    if (r <= 5.0) {
        const double a = r - 1.6;
        m = PE_AS241_C(double, a) / PE_AS241_D(double, a);
    } else {
        const double a = r - 5.0;
        m = PE_AS241_E(double, a) / PE_AS241_F(double, a);
    }
    return centered < 0.0 ? -m : m;
}

// The same FP64 arithmetic for kernels with many other live values: every
// coefficient passes through an empty asm statement, so it is materialised
// in scalar registers at its use instead of being hoisted out of the
// kernel's loops into 50-odd vector-register pairs.  Bit-identical results.
__device__ __forceinline__ double pe_sgpr(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(c));
#endif
    return c;
}

__device__ __forceinline__ double normal_from_word_f64_lean(uint32_t word) {
    const double p = (static_cast<double>(word) + 0.5) * (1.0 / 4294967296.0);
    const double centered = p - 0.5;
    if (fabs(centered) <= 0.425) {
        const double a = pe_sgpr(0.180625) - centered * centered;
        return centered * PE_AS241_A(pe_sgpr, a) / PE_AS241_B(pe_sgpr, a);
    }
    const double tail = centered < 0.0 ? p : 1.0 - p;
    const double r = sqrt(-log(tail));
    double m;
    if (r <= 5.0) {
        const double a = r - 1.6;
        m = PE_AS241_C(pe_sgpr, a) / PE_AS241_D(pe_sgpr, a);
    } else {
        const double a = r - 5.0;
        m = PE_AS241_E(pe_sgpr, a) / PE_AS241_F(pe_sgpr, a);
    }
    return centered < 0.0 ? -m : m;
}

// FP32: the same rational approximations evaluated in FP32.  The centred
// coordinate and the tail probability come from the integer word directly,
// so FP32 rounding never maps a word to 0, 1 or the wrong side of 1/2:
//   p - 1/2 = (int32(w ^ 2^31) + 1/2) 2^-32,  tail = (min(w, ~w) + 1/2) 2^-32.
PE_HD float normal_from_word_f32(uint32_t word) {
    const float centered = (static_cast<float>(static_cast<int32_t>(word ^ 0x80000000u)) + 0.5f) * 2.3283064365386963e-10f;
    if (fabsf(centered) <= 0.425f) {
        const float a = 0.180625f - centered * centered;
        return centered * PE_AS241_A(float, a) / PE_AS241_B(float, a);
    }
    const uint32_t low = word < 0x80000000u ? word : ~word;
    const float tail = (static_cast<float>(low) + 0.5f) * 2.3283064365386963e-10f;
    const float r = sqrtf(-logf(tail));
    float m;
    if (r <= 5.0f) {
        const float a = r - 1.6f;
        m = PE_AS241_C(float, a) / PE_AS241_D(float, a);
    } else {
        const float a = r - 5.0f;
        m = PE_AS241_E(float, a) / PE_AS241_F(float, a);
    }
    return centered < 0.0f ? -m : m;
}

template <typename Real>
PE_HD Real normal_from_word(uint32_t w);
template <>
PE_HD double normal_from_word<double>(uint32_t w) { return normal_from_word_f64(w); }
template <>
PE_HD float normal_from_word<float>(uint32_t w) { return normal_from_word_f32(w); }

// ---------------------------------------------------------- accumulators --
// Per-lane statistics of one component over the units a lane processes:
// Neumaier sum + Welford moments (reduction.rs:3-161).
struct LaneStat {
    double sum, corr, mean, m2;
    uint32_t n;

    PE_HD void init() { sum = corr = mean = m2 = 0.0; n = 0; }
    PE_HD void add(double v) {
        const double t = sum + v;
        if (fabs(sum) >= fabs(v)) corr += (sum - t) + v; else corr += (v - t) + sum;
        sum = t;
        const uint32_t nn = n + 1;
        const double d = v - mean;
        const double nm = mean + d / static_cast<double>(nn);
        m2 += d * (v - nm);
        mean = nm;
        n = nn;
    }
    // Chan-Golub-LeVeque merge of moments and a Neumaier merge of sums.
    PE_HD void merge(const LaneStat& o) {
        if (o.n == 0) return;
        if (n == 0) { *this = o; return; }
        double parts[3] = {corr, o.sum, o.corr};
        double s = sum, c = 0.0;
        for (int i = 0; i < 3; ++i) {
            const double v = parts[i];
            const double t = s + v;
            if (fabs(s) >= fabs(v)) c += (s - t) + v; else c += (v - t) + s;
            s = t;
        }
        sum = s;
        corr = c;
        const uint32_t nn = n + o.n;
        const double delta = o.mean - mean;
        const double lw = static_cast<double>(n), rw = static_cast<double>(o.n), tw = static_cast<double>(nn);
        mean = mean + delta * (rw / tw);
        m2 = (m2 + o.m2) + delta * delta * (lw * rw / tw);
        n = nn;
    }
};

// Block partial written to HBM: one per (block, component).
struct BlockStat {
    double sum, corr, mean, m2;
    uint64_t n;
};

}  // namespace pe
