// exact_div.hh -- correctly rounded division of a double by 3 without a
// division instruction.  Plain C++ on the host (std::fma), HIP on the device,
// so the host-side check runs the very code the kernels use.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define HONEI_HD __host__ __device__
#else
#define HONEI_HD
#endif

namespace honei {
namespace hipx {

// x / 3 correctly rounded (what the division x / 3.0 returns) with one
// multiply and two fused multiply-adds instead of a full division:
// q0 = x * RN(1/3) is within 1.25 ulp of x/3; e = x - 3 q0 is a multiple of
// that ulp of at most 3.75 ulp, so the fma forms it exactly; q0 + e RN(1/3)
// lies within 2^-54 ulp of x/3, and x/3 (x is a multiple of two ulps of x/3,
// so x/3 is k/3 ulp) stays at least ulp/6 from every rounding midpoint, hence
// rounding q0 + e RN(1/3) once gives RN(x/3).  Normal, finite x (the cell
// coefficients and their sums).  x / 6 = div3_rn(x) * 0.5 exactly.
HONEI_HD inline double div3_rn(double x) {
    const double r3 = 1.0 / 3.0;
    const double q0 = x * r3;
    const double e = std::fma(-q0, 3.0, x);
    return std::fma(e, r3, q0);
}

}  // namespace hipx
}  // namespace honei
