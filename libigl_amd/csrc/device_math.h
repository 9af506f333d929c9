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

// FP32 far field of an order-0/1/2 Taylor expansion written as one
// polynomial in the direction to the query.  For the offset q (l2 = |q|^2),
// x = q/|q| (kept normalised so that the cubic cannot overflow for distant
// queries) and m = 1/|q| the value is
//   m^2 [ x.b + m ( Qf(x) + m G(x) ) ]
// with b = b.xyz, the quadratic form Qf(x) = c.x x^2 + c.y y^2 + c.z z^2 +
// c.w xy + d.x yz + d.y xz and the cubic G(x) = d.z x^3 + d.w y^3 + e.x z^3 +
// e.y x^2y + e.z x^2z + e.w xy^2 + f.x y^2z + f.y xz^2 + f.z yz^2 + f.w xyz.
// The soup's slot records (k_write_slots) and the cloud's cell records
// (k_rec_poly) store their expansions in this form.
template <int ORDER>
__device__ __forceinline__ float far_field_poly(float x, float y, float z, float l2, const float4& b,
                                                const float4& c, const float4& d, const float4& e,
                                                const float4& f) {
  // explicit multiply-adds (no compiler contraction): every kernel that
  // inlines this computes each term bit for bit the same
  const float m = __builtin_amdgcn_rsqf(l2);
  x *= m;
  y *= m;
  z *= m;
  float t = fmaf(z, b.z, fmaf(y, b.y, x * b.x));  // x.b
  if (ORDER >= 1) {
    const float z2 = z * z;
    const float q1 = fmaf(d.y, z, fmaf(c.w, y, c.x * x)), q2 = fmaf(d.x, z, c.y * y);
    float qf = fmaf(c.z, z2, fmaf(y, q2, x * q1));
    if (ORDER >= 2) {
      const float g1 = fmaf(e.z, z, fmaf(e.y, y, d.z * x)), g2 = fmaf(f.w, z, e.w * y);
      const float gx = fmaf(f.y, z2, fmaf(y, g2, x * g1));
      const float gy = fmaf(f.z, z2, y * fmaf(f.x, z, d.w * y));
      const float g = fmaf(e.x, z2 * z, fmaf(y, gy, x * gx));
      qf = fmaf(m, g, qf);
    }
    t = fmaf(m, qf, t);
  }
  return (m * m) * t;
}

}  // namespace gwn
