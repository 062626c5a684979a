// Device helpers shared by the far-field kernels (HIP sources only).
#pragma once

#include <hip/hip_runtime.h>

namespace sphstk {

// FP64 traction: the Newton correction of v_rsq_f64 (y0 = r^-1 (1 + delta),
// |delta| <= 5.3e-8) fused into the fifth power.  r^-5 = y0^5 (1 - 5 delta +
// O(delta^2)) and 1 - 5 delta = 3.5 - 2.5 r^2 y0^2 + O(delta^2)
// = -2.5 (r^2 y0^2 - 1.4) + O(delta^2), so
//   w = (y0^2)^2 * (y0 * fma(r^2, y0^2, -1.4)) = -0.4 r^-5 (1 + O(17.5 delta^2))
// and the kernel applies the factor kTracW = -2.5 once per target.
constexpr double kTracW = -2.5;
__device__ __forceinline__ double trac_w_f64(double r2) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double t = y0 * y0;
    return (t * t) * (y0 * __builtin_fma(r2, t, -1.4));
}

} // namespace sphstk
