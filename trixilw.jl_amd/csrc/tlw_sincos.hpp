// tlw_sincos.hpp -- sin and cos of a moderate double argument for the source
// terms, without private memory.
//
// Cody-Waite reduction by π/2 (W. J. Cody and W. Waite, Software Manual for the
// Elementary Functions, 1980): k = rint(a · 2/π), r = ((a - k P1) - k P2) - k P3
// with π/2 = P1 + P2 + P3 + O(2^-114), P1 and P2 holding 30 significant bits
// each, so that k P1 and k P2 are exact for |k| < 2^23 (|a| < 6.5e6); then the
// Taylor series of sin and cos at |r| <= π/4 (terms to r^19 and r^20, whose
// next terms are below 2^-60 of the result there), and the quadrant k mod 4.
// Accurate to about 1 ulp on that range; the source terms evaluate it at
// |a| = π|x + y - t| of the domain and time.
#pragma once

namespace tlw {

__device__ __forceinline__ void sincos_cw(double a, double* s, double* c) {
    constexpr double kTwoOverPi = 0.6366197723675814;
    constexpr double P1 = 1.570796325802803;       // π/2 to 30 bits
    constexpr double P2 = 9.920935791635221e-10;   // the next 30 bits
    constexpr double P3 = 5.170182981794105e-19;   // the rest
    const double k = rint(a * kTwoOverPi);
    const double r = ((a - k * P1) - k * P2) - k * P3;
    const double r2 = r * r;
    // sin r = r (1 - r²/3! + r⁴/5! - ... - r^18/19!), cos r = 1 - r²/2! + ... + r^20/20!
    double ps = -1.0 / 121645100408832000.0;  // -1/19!
    ps = fma(ps, r2, 1.0 / 355687428096000.0);
    ps = fma(ps, r2, -1.0 / 1307674368000.0);
    ps = fma(ps, r2, 1.0 / 6227020800.0);
    ps = fma(ps, r2, -1.0 / 39916800.0);
    ps = fma(ps, r2, 1.0 / 362880.0);
    ps = fma(ps, r2, -1.0 / 5040.0);
    ps = fma(ps, r2, 1.0 / 120.0);
    ps = fma(ps, r2, -1.0 / 6.0);
    const double sr = fma(r * r2, ps, r);
    double pc = 1.0 / 2432902008176640000.0;  // 1/20!
    pc = fma(pc, r2, -1.0 / 6402373705728000.0);
    pc = fma(pc, r2, 1.0 / 20922789888000.0);
    pc = fma(pc, r2, -1.0 / 87178291200.0);
    pc = fma(pc, r2, 1.0 / 479001600.0);
    pc = fma(pc, r2, -1.0 / 3628800.0);
    pc = fma(pc, r2, 1.0 / 40320.0);
    pc = fma(pc, r2, -1.0 / 720.0);
    pc = fma(pc, r2, 1.0 / 24.0);
    pc = fma(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    const int q = (int)(k - 4.0 * floor(k * 0.25));  // k mod 4 in 0..3
    const double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

}  // namespace tlw
