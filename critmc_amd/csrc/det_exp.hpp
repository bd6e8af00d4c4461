// e^x as a fixed sequence of IEEE double operations: Cody-Waite reduction
// x = n ln2 + r and the degree-13 Taylor polynomial of e^r in Horner / fma
// form, scaled by 2^n.  Any correct IEEE implementation of these steps gives
// the same bits, so accept / reject decisions that compare a uniform with it
// are reproducible on the host.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace cm {

__host__ __device__ inline double det_exp(double x)
{
    if (x < -745.0) return 0.0;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double n = rint(x * inv_ln2);
    double r = fma(-n, ln2_hi, x);
    r = fma(-n, ln2_lo, r);
    double p = 1.0 / 6227020800.0;
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

}  // namespace cm
