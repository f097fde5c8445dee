// Log-sum-exp accumulators (ac::Lse, value = m + log1p(t)) shared by the fused kernels and the stage operators.
#pragma once

#include <hip/hip_runtime.h>

#include "ac_internal.hpp"

namespace ac {

// One exp per value and no branches: the exponent is always min - max of (x, m), which rescales t when x is a new
// maximum and is the new term otherwise.  exp(-inf) = 0 covers the empty accumulator.  min - max is formed as
// -|x - m|: rounding to nearest is sign-symmetric, so fl(min - max) = -|fl(x - m)| exactly, and x > m exactly when
// fl(x - m) > 0 (the difference of two distinct finite doubles never rounds to zero; m = -inf gives +inf).
__device__ __forceinline__ void lse_add(Lse& s, double x) {
  const bool valid = x != -__builtin_huge_val();
  const double d = x - s.m;  // NaN only for x = m = -inf, which is not valid
  const double e = (double)__expf((float)(-fabs(d)));
  const double t = d > 0.0 ? (1.0 + s.t) * e : s.t + e;
  s.t = valid ? t : s.t;
  s.m = valid ? fmax(x, s.m) : s.m;
}

// lse_add for a finite x (no -inf term to skip): the same result with fewer selects.  An empty accumulator
// (m = -inf) takes x as its maximum: exp(-inf) = 0 clears t.
__device__ __forceinline__ void lse_add_finite(Lse& s, double x) {
  const double d = x - s.m;
  const double e = (double)__expf((float)(-fabs(d)));
  s.t = d > 0.0 ? (1.0 + s.t) * e : s.t + e;
  s.m = fmax(x, s.m);
}

__device__ __forceinline__ Lse lse_merge(Lse a, Lse b) {
  if (b.m == -__builtin_huge_val()) return a;
  if (a.m == -__builtin_huge_val()) return b;
  if (a.m >= b.m) {
    a.t += (1.0 + b.t) * exp(b.m - a.m);
    return a;
  }
  b.t += (1.0 + a.t) * exp(a.m - b.m);
  return b;
}

// lse_merge with the exponential of the two maxima's difference in FP32 (__expf), as lse_add forms each term: for merges
// in a kernel's inner loop, where the FP64 exp would dominate.  The merged t keeps a relative error of about 1e-7.
__device__ __forceinline__ Lse lse_merge_fast(Lse a, Lse b) {
  if (b.m == -__builtin_huge_val()) return a;
  if (a.m == -__builtin_huge_val()) return b;
  if (a.m >= b.m) {
    a.t += (1.0 + b.t) * (double)__expf((float)(b.m - a.m));
    return a;
  }
  b.t += (1.0 + a.t) * (double)__expf((float)(a.m - b.m));
  return b;
}

__device__ __forceinline__ Lse lse_empty() { return Lse{-__builtin_huge_val(), 0.0}; }

}  // namespace ac
