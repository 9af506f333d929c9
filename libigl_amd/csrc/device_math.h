// device_math.h -- FP32 device helpers shared by the traversal kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace gwn {

// atan on [0,1] as a * P(a^2), P of degree 8: a weighted least-squares /
// Lawson fit to atan(x)/x on 4000 Chebyshev nodes (computed for this
// library); max error 1.2e-8 in exact arithmetic, ~1e-7 rad when evaluated in
// FP32 (comparable to a correctly rounded atan2f).
__device__ __forceinline__ float atan01(float a) {
  const float s = a * a;
  float p = 0.0028501534834504128f;
  p = fmaf(p, s, -0.016069747507572174f);
  p = fmaf(p, s, 0.04269346594810486f);
  p = fmaf(p, s, -0.07504474371671677f);
  p = fmaf(p, s, 0.10641028732061386f);
  p = fmaf(p, s, -0.14203672111034393f);
  p = fmaf(p, s, 0.1999262422323227f);
  p = fmaf(p, s, -0.33333075046539307f);
  p = fmaf(p, s, 1.0f);
  return a * p;
}

// atan2 for y != 0 (the caller returns early on a zero numerator).
__device__ __forceinline__ float atan2_fast(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  float r = atan01(mn * __builtin_amdgcn_rcpf(mx));
  if (ay > ax) r = 1.57079632679489662f - r;
  if (x < 0.0f) r = 3.14159265358979324f - r;
  return copysignf(r, y);
}

}  // namespace gwn
