// angle(a./b) of the reference's winding and plaquette sums, on complex doubles.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace hqt {

// angle(a/b) = atan2 of a*conj(b) for b != 0 (same angle; no division).  Exact zeros follow
// IEEE component division, as MATLAB's and NumPy's a./b do: a == 0 gives 0, and b == 0
// divides each component of a by +0 (a/0 is then +-Inf or NaN per component, and the angle is
// a quadrant diagonal or NaN; 0/0 is NaN).
__device__ __forceinline__ double angle_ratio(double ar, double ai, double br, double bi) {
    if (ar == 0.0 && ai == 0.0) return (br == 0.0 && bi == 0.0) ? __builtin_nan("") : 0.0;
    if (br == 0.0 && bi == 0.0) return atan2(ai / 0.0, ar / 0.0);
    return atan2(ai * br - ar * bi, ar * br + ai * bi);
}

}  // namespace hqt
