// exp(-x/2) in FP64 for x >= 0: the common weight of the weighted Laguerre
// functions (x = S / K), shared by the device code (pe_lsm.inc) and a CPU
// accuracy check against the host's extended-precision exponential.
//
// a = -x/2 is clamped at -1000 (the result underflows to zero there either
// way) and split as a = n ln 2 + r with n = rint(a / ln 2), |r| <= ln(2)/2;
// the product n ln 2 is subtracted in two parts (Cody and Waite, "Software
// Manual for the Elementary Functions", 1980): ln2_hi keeps 32 significant
// bits, so n ln2_hi is exact for |n| < 2^21.  exp(r) is its Taylor series to
// degree 13 in Horner form (the first omitted term is below 6e-18 relative on
// |r| <= 0.347), scaled by 2^n.  About 20 instructions on the device, with
// no special-case branches: the argument range needs none.
#pragma once

#include <math.h>

#ifndef PE_NM_HD
#define PE_NM_HD inline
#endif

namespace pe {

// fma(a, b, c) with the constant c in a scalar register pair: one
// v_fma_f64, where the compiler's two-address form (v_fmac_f64 with the
// constant copied into the destination first) would add a 64-bit move per
// step and hold the constants in vector registers.
PE_NM_HD double fma_sc(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
#else
    return fma(a, b, c);
#endif
}

PE_NM_HD double exp_neg_half(double x) {
    const double a = fmax(-0.5 * x, -1000.0);
    const double n = rint(a * 1.4426950408889634);
    double r = fma(-n, 6.9314718036912382e-01, a);
    r = fma(-n, 1.9082149292705877e-10, r);
    double p = 1.6059043836821613e-10;   // 1/13!
    p = fma_sc(p, r, 2.0876756987868100e-09);
    p = fma_sc(p, r, 2.5052108385441720e-08);
    p = fma_sc(p, r, 2.7557319223985890e-07);
    p = fma_sc(p, r, 2.7557319223985893e-06);
    p = fma_sc(p, r, 2.4801587301587302e-05);
    p = fma_sc(p, r, 1.9841269841269841e-04);
    p = fma_sc(p, r, 1.3888888888888889e-03);
    p = fma_sc(p, r, 8.3333333333333332e-03);
    p = fma_sc(p, r, 4.1666666666666664e-02);
    p = fma_sc(p, r, 1.6666666666666666e-01);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, static_cast<int>(n));
}

}  // namespace pe
