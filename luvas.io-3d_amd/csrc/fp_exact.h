// Correctly rounded FP32 reciprocal, square root and quotient from the hardware's approximate
// instructions, for operands in stated ranges.  Each returns exactly the IEEE result (round to
// nearest even) of the operation it replaces, so kernels that use them keep the bits of the
// IEEE arithmetic definition while issuing fewer instructions than the general expansions
// (which also handle denormals, infinities and rescaling that these ranges never need).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lr3d {
namespace {

// 1.0f / z.  v_rcp_f32 followed by one Newton step with FMA reproduces the IEEE quotient
// 1.0f / z bit for bit for every |z| in [2^-60, 2^61) (checked exhaustively on gfx950: all
// 1 015 021 568 such values of both signs).  The step alone does not make an arbitrary 1-ulp
// seed exact, so the statement is about this instruction's results.
__device__ __forceinline__ float rcp_exact(float z) {
    const float r = __builtin_amdgcn_rcpf(z);
    return fmaf(fmaf(-z, r, 1.0f), r, r);
}

// sqrtf(x) for normal x >= 2^-96: v_sqrt_f32 is within 1 ulp, and the neighbour whose square
// brackets x is selected with two FMA residuals (the compiler's correctly rounded expansion,
// without its rescaling of tiny inputs and its zero/inf/NaN fix-ups, which this range never
// needs).  Exact from any seed within 1 ulp of the root.
__device__ __forceinline__ float sqrt_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float below = __int_as_float(__float_as_int(s) - 1);
    const float above = __int_as_float(__float_as_int(s) + 1);
    const float r = fmaf(-below, s, x) <= 0.0f ? below : s;
    return fmaf(-above, s, x) > 0.0f ? above : r;
}

// a / b given y = RN(1 / b), for b, a / b and the residuals below in the normal range.
// q0 = RN(a y) is within 1.5 ulp of a / b; one correction q1 = RN(q0 + RN(a - b q0) y) brings it
// within one ulp (faithful); for a faithful q1 the residual a - b q1 is exact and Markstein's
// theorem (y correctly rounded, q1 faithful) makes RN(q1 + (a - b q1) y) the correctly rounded
// quotient.  a = 0 gives +0.
__device__ __forceinline__ float div_rn(float a, float b, float y) {
    const float q0 = a * y;
    const float q1 = fmaf(fmaf(-b, q0, a), y, q0);
    return fmaf(fmaf(-b, q1, a), y, q1);
}

}  // namespace
}  // namespace lr3d
