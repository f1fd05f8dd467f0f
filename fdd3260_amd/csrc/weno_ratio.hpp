// weno_ratio.hpp -- the smoothness-ratio switch of the reference 'weno'
// weights (advq_weno.f90:339-349: nonlinear weights when
// max(eta)/(min(eta)+eps) > 20) evaluated without the division but with the
// same outcome as the correctly rounded quotient.  Shared by the plain FP64
// forward kernel and the adjoint, so that both take the branch the CPU
// restatement takes.
#pragma once
#include <hip/hip_runtime.h>

namespace mimadv {

// fl(x/d) > 20 without the division, d > 0: the quotient rounds above 20
// iff x/d > 20 + 2^-49 (that tie rounds to the even 20), i.e. iff
// x - 20d > h = 2^-49 d.  y = fl(x - 20d) from one fma has relative error at
// most 2^-53 and h is exact, so outside a 2^-40 band around h the answer is
// certain; inside it (never met in practice) the division decides.
__device__ __forceinline__ bool ratio_above_20(double x, double d)
{
    const double y = fma(-20.0, d, x);
    const double h = ldexp(d, -49);
    if (y > h * (1.0 + 0x1p-40)) return true;
    if (y < h * (1.0 - 0x1p-40)) return false;
    return x / d > 20.0;
}

}  // namespace mimadv
