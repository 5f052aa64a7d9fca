// Correctly rounded FP32 division by a constant through its reciprocal and one
// fused correction (P. Markstein, "Computation of elementary functions on the
// IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990; the division
// theorem as stated in Muller et al., Handbook of Floating-Point Arithmetic,
// ch. 4.7): with y = RN(1/b), q0 = RN(a y) lies within one ulp of a/b, the
// remainder r = b q0 - a is exact under an fma, and RN(q0 - r y) is a/b
// rounded to nearest, provided the intermediate values stay normal.
//
// The fused sweep divides |p| by alpha1 and |q| by alpha0 this way (three
// instructions instead of the IEEE operator's ten), every voxel and every
// iteration.  Bounds: b within [2^-60, 2^36] (the host takes the fused sweep
// only then); a finite and >= 0.  For a >= 2^-90 every intermediate is normal
// and the quotient is exact; for 0 < a < 2^-90 it may be off in its last bit,
// but then a/b < 2^-30 and the sweep only compares it with 1; a = 0 gives +0.
// Checked on the host against double-precision division (its double rounding
// to 24 bits is innocuous, 53 >= 2 * 24 + 2) over random, near-midpoint and
// exact-quotient numerators, on the device against the IEEE operator, and by
// the bitwise parity tests of the fused sweep against the two-sweep loop and
// the oracle, which divide with the IEEE operator.
#pragma once
#include <hip/hip_runtime.h>

namespace amb {

// RN(1/alpha0), RN(1/alpha1) for a fused-sweep launch.
struct DivConsts {
    float inv_alpha0, inv_alpha1;
};

inline bool div_rn_divisor_ok(float b) { return b >= 0x1p-60f && b <= 0x1p36f; }

inline DivConsts div_consts(float alpha0, float alpha1) {
    // double then float: RN(1/b), the double rounding being innocuous
    return DivConsts{float(1.0 / double(alpha0)), float(1.0 / double(alpha1))};
}

// a / b rounded to nearest, given y = RN(1/b), within the bounds above.
__device__ __forceinline__ float div_rn(float a, float b, float y) {
    const float q0 = a * y;
    const float r = fmaf(b, q0, -a);
    return fmaf(-r, y, q0);
}

}  // namespace amb
