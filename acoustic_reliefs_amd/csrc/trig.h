// FP64 sine and cosine of one argument for the Green's-function kernels, host and device.
//
// Method: Cody-Waite argument reduction (W. J. Cody, W. Waite, "Software Manual for the Elementary
// Functions", 1980) with a three-part pi/2 and fused multiply-adds, x = n pi/2 + r, |r| <= pi/4 (a
// little more when n is rounded across a quadrant boundary), then the Taylor series of sin r and
// cos r evaluated by Horner's rule, truncated where the remainder is below 1e-19 on |r| <= pi/4 + 1e-3
// (r^19 / 19! and r^20 / 20!), and the quadrant n mod 4 selecting and negating the two.
//
// Domain: |x| <= kTrigMax. The three-part pi/2 leaves n * 5.6e-50 unrepresented and each reduction
// step rounds once, so for |x| up to kTrigMax the absolute error of both results is a few 1e-16
// (tests/test_trig.py measures it against the C library). The kernels' arguments are k r, k (s . c)
// with distances of metres and k of at most a few hundred per metre, far inside the domain; callers
// check the wavenumber against it (kTrigMaxK).
//
// Every operation is an explicit fma or a plain product, so the host and device builds round alike.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define ARH_HD __host__ __device__
#else
#define ARH_HD
#endif

namespace arh {

constexpr double kTrigMax = 1.0e9;  // largest |x| the reduction is accurate for
constexpr double kTrigMaxK = 1.0e6;  // wavenumber limit the entry points enforce (k r <= kTrigMax up to 1 km)

ARH_HD inline void trig_sincos(double x, double *s, double *c) {
    // pi/2 = P1 + P2 + P3 (+ 5.6e-50), each the double nearest to what is left
    const double kTwoOverPi = 0.6366197723675814;
    const double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
    const double n = rint(x * kTwoOverPi);
    double r = fma(-n, P1, x);
    r = fma(-n, P2, r);
    r = fma(-n, P3, r);
    const double r2 = r * r;
    // sin r = r + r^3 (-1/3! + r^2 (1/5! - r^2 (...))), up to r^17 / 17!
    double ps = 1.0 / 355687428096000.0;      // 1/17!
    ps = fma(ps, r2, -1.0 / 1307674368000.0);  // 1/15!
    ps = fma(ps, r2, 1.0 / 6227020800.0);      // 1/13!
    ps = fma(ps, r2, -1.0 / 39916800.0);       // 1/11!
    ps = fma(ps, r2, 1.0 / 362880.0);          // 1/9!
    ps = fma(ps, r2, -1.0 / 5040.0);           // 1/7!
    ps = fma(ps, r2, 1.0 / 120.0);             // 1/5!
    ps = fma(ps, r2, -1.0 / 6.0);              // 1/3!
    const double sr = fma(r * r2, ps, r);
    // cos r = 1 + r^2 (-1/2! + r^2 (1/4! - r^2 (...))), up to r^18 / 18!
    double pc = -1.0 / 6402373705728000.0;     // 1/18!
    pc = fma(pc, r2, 1.0 / 20922789888000.0);  // 1/16!
    pc = fma(pc, r2, -1.0 / 87178291200.0);    // 1/14!
    pc = fma(pc, r2, 1.0 / 479001600.0);       // 1/12!
    pc = fma(pc, r2, -1.0 / 3628800.0);        // 1/10!
    pc = fma(pc, r2, 1.0 / 40320.0);           // 1/8!
    pc = fma(pc, r2, -1.0 / 720.0);            // 1/6!
    pc = fma(pc, r2, 1.0 / 24.0);              // 1/4!
    pc = fma(pc, r2, -0.5);                    // 1/2!
    const double cr = fma(pc, r2, 1.0);
    // quadrant: x = r + q pi/2 (mod 2 pi)
    const int q = (int)((long long)n & 3);
    const double a = (q & 1) ? cr : sr;  // |sin x|'s source
    const double b = (q & 1) ? sr : cr;  // |cos x|'s source
    *s = (q & 2) ? -a : a;
    *c = ((q + 1) & 2) ? -b : b;
}

}  // namespace arh
