// Inverse normal of the mixed-precision kernels, as plain C++ shared by the
// device code (pe_kernels.inc, native v_log_f32 / v_sqrt_f32) and a CPU
// accuracy check over every 32-bit word.
//
// z = sqrt(2) erfinv(x), x = 2u - 1, with erfinv(x) = x p(w) and
// w = -log(1 - x^2) = -log(4 s (1 - s)), where s = min(u, 1 - u) comes from
// the integer word so both tails keep full precision.  For w < 5 (|z| <
// 2.93: all but 0.34% of the draws) p is a degree-9 polynomial in w - 2.5;
// beyond, a degree-9 polynomial in sqrt(w) - 3.  The coefficients are
// weighted least-squares fits of erfinv(x)/x over those ranges made for this
// engine (the form of Giles, "Approximating the erfinv function", GPU
// Computing Gems, 2011).
#pragma once

#include <stdint.h>

#ifndef PE_NM_HD
#define PE_NM_HD inline
#endif

namespace pe {

// 2^32 (p - 1/2) and s = min(p, 1 - p) of the word's midpoint probability,
// exact in integer arithmetic before one FP32 rounding (no word maps to 0, 1
// or the wrong side of 1/2).  The centred value is left scaled by 2^32: the
// power of two is folded into nm_central's constant, which rounds the same.
PE_NM_HD float nm_centered(uint32_t word) {
    return static_cast<float>(static_cast<int32_t>(word ^ 0x80000000u)) + 0.5f;
}

PE_NM_HD float nm_tail_probability(uint32_t word) {
    const uint32_t low = word < 0x80000000u ? word : ~word;
    return (static_cast<float>(low) + 0.5f) * 2.3283064365386963e-10f;
}

// w < 5: z = 2 sqrt(2) (p - 1/2) p(w - 2.5), with centered = 2^32 (p - 1/2)
PE_NM_HD float nm_central(float centered, float w) {
    const float v = w - 2.5f;
    float p = -9.315122007e-09f;
    p = p * v + 2.280885703e-08f;
    p = p * v + 4.730939906e-07f;
    p = p * v - 3.457865205e-06f;
    p = p * v - 4.993500625e-06f;
    p = p * v + 2.183274482e-04f;
    p = p * v - 1.252694055e-03f;
    p = p * v - 4.177368246e-03f;
    p = p * v + 2.466402650e-01f;
    p = p * v + 1.501409411e+00f;
    return ((2.8284271247461903f * 2.3283064365386963e-10f) * centered) * p;
}

// w >= 5: z = sqrt(2) (1 - 2 s) p(sqrt(w) - 3), signed by the word's half
PE_NM_HD float nm_tail(uint32_t word, float s, float sqrt_w) {
    const float v = sqrt_w - 3.0f;
    float p = 3.044116056e-05f;
    p = p * v - 1.220943814e-04f;
    p = p * v - 6.469927757e-05f;
    p = p * v + 1.277914853e-03f;
    p = p * v - 3.523040097e-03f;
    p = p * v + 5.760414060e-03f;
    p = p * v - 7.665638812e-03f;
    p = p * v + 9.436299093e-03f;
    p = p * v + 1.001677275e+00f;
    p = p * v + 2.832977057e+00f;
    const float z = (1.4142135623730951f * (1.0f - 2.0f * s)) * p;
    return word < 0x80000000u ? -z : z;
}

}  // namespace pe
