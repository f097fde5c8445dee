// Correction step of a correctly rounded FP32 square root, shared by the distance kernels and their CPU check.
#pragma once
#include <hip/hip_runtime.h>

namespace ac {

// Given x >= 0 and s within 1 ulp of sqrt(x) (the hardware root), the float nearest sqrt(x): s or one of its
// neighbours s -/+ 1 ulp, chosen by the signs of the fused residuals x - s' s.  x at or below the product of s and its
// lower neighbour puts the root at or below their midpoint, x above the product with the upper neighbour puts it
// above theirs.  Exact for every x >= 2^-104 (tests/sqrt_rn_check.hip checks it over the floats); below, the
// residuals lose bits to underflow and the result can be one ulp off (distances under 1e-16).
__host__ __device__ inline float sqrt_rn_fix(float x, float s) {
  const unsigned b = __builtin_bit_cast(unsigned, s);
  const float sd = __builtin_bit_cast(float, b - 1u), su = __builtin_bit_cast(float, b + 1u);
  float r = s;
  if (fmaf(-sd, s, x) <= 0.f) r = sd;
  if (fmaf(-su, s, x) > 0.f) r = su;
  return r;
}

}  // namespace ac
