// lsh_math.h -- float64 elementary functions of the march and the lighting
// path, written for their inputs (finite, moderate magnitudes), shared by the
// HIP kernels and the host builds (no HIP headers: the tests compile it with a
// host compiler), and the wave-vote helpers their fast paths use.
#pragma once

#if defined(__HIPCC__)
#define LSH_MHD __host__ __device__ inline
#else
#define LSH_MHD inline
#endif

namespace lsh {

// True in every lane when the predicate holds in any active lane of the
// wavefront (a wave-uniform branch condition); one lane on the host.
LSH_MHD bool lsh_any_lane(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ballot_w64(p) != 0;
#else
  return p;
#endif
}

// A wave-uniform integer as the compiler's scalar value (read from the first
// active lane), so that branches on it are scalar branches; x on the host.
LSH_MHD int lsh_uniform(int x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_readfirstlane(x);
#else
  return x;
#endif
}

// An empty statement the compiler may not move or remove: placed in the rare
// side of a wave-uniform branch, it keeps the branch a branch (the compiler
// would otherwise evaluate both sides for every lane and select).
LSH_MHD void lsh_keep_branch() {
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__ volatile("" ::: "memory");
#endif
}

// The value unchanged, but opaque to the compiler (no instruction): tests formed
// from it are not merged with tests of the original.
template <class T>
LSH_MHD void lsh_keep_value(T& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__ volatile("" : "+v"(x));
#else
  (void)x;
#endif
}

// Correctly rounded double square root and quotient for the operands the march
// and the lighting form. On the device these are the steps of the compiler's IEEE sequences --
// reciprocal-square-root or reciprocal estimate, Goldschmidt / Newton
// refinement, a final residual correction -- without the operand scaling those
// sequences apply only when an exponent nears the ends of the double range (a
// radicand below 2^-767; a quotient whose operands lie 768 binades apart, or
// whose numerator is below 2^-969 or zero) and without the special-value
// fix-up (zero, infinity, NaN). For the operands below the results equal the
// host's sqrt and division bit for bit.
//
// lsh_sqrt_normal: finite x >= 2^-767.
LSH_MHD double lsh_sqrt_normal(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
#else
  return __builtin_sqrt(x);
#endif
}
// sqrt(x) for any x >= 0: the sequence above unless a lane of the wavefront
// holds a radicand below 2^-767 (then the general one, for every lane).
LSH_MHD double lsh_sqrt_any(double x) {
  if (lsh_any_lane(!(x >= 0x1p-767))) return __builtin_sqrt(x);
  return lsh_sqrt_normal(x);
}
// lsh_quotient: a / b for |a| < 2^1000 (a = 0 or |a| >= 2^-969) and
// 2^-200 < b < 2^200, as the march's cell exits and the arctangent's reduced
// argument are. A zero numerator gives +0.
LSH_MHD double lsh_quotient(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = a * r;
  e = __builtin_fma(-b, q, a);
  return __builtin_fma(e, r, q);
#else
  return a / b;
#endif
}
// a / b with the sign of zero and every special operand as IEEE division (the
// sequence's final fix-up kept), for normal b whose quotient with a stays
// normal or zero: the slopes and step lengths of the march.
LSH_MHD double lsh_ratio(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_div_fixup(lsh_quotient(a, b), b, a);
#else
  return a / b;
#endif
}
LSH_MHD float lsh_quotient(float a, float b) { return a / b; }
LSH_MHD float lsh_ratio(float a, float b) { return a / b; }
LSH_MHD float lsh_sqrt_any(float x) { return __builtin_sqrtf(x); }


// Float64 arctangent of y / x in (-pi, pi] for finite (y, x), not both zero
// (the pixel-to-body vectors of the lighting), by the textbook reductions
//   atan(y / x) = pi/2 - atan(x / y) when |y| > |x|, and
//   atan(t) = pi/4 + atan((t - 1) / (t + 1)) when t > tan(pi/8),
// taken as one division by selecting its numerator and denominator, then
// atan(q) = q + q^3 P(q^2) on |q| <= tan(pi/8) with the degree-10 polynomial P
// of atan_coefficients.py (|error| <= 3e-18 there), and the quadrant from the
// signs of x and y. The general library routine also handles infinities,
// signed zeros and NaN, which these inputs never hold, at several times the
// instruction count.
LSH_MHD double lsh_fma64(double a, double b, double c) { return __builtin_fma(a, b, c); }

LSH_MHD double atan2_finite(double y, double x) {
  constexpr double kTanPi8 = 0.41421356237309504880;  // sqrt(2) - 1
  constexpr double kPi4 = 0.78539816339744830962, kPi2 = 1.57079632679489661923;
  constexpr double kPiHi = 3.14159265358979323846;
  const double ax = x < 0.0 ? -x : x, ay = y < 0.0 ? -y : y;
  const bool swap = ay > ax;
  const double num = swap ? ax : ay, den = swap ? ay : ax;  // num / den in [0, 1]
  const bool upper = num > kTanPi8 * den;
  const double q = lsh_quotient(upper ? num - den : num, upper ? num + den : den);
  const double z = q * q;
  double p = -0.01917688711906226;
  p = lsh_fma64(p, z, 0.03923165829558719);
  p = lsh_fma64(p, z, -0.0508544973794026);
  p = lsh_fma64(p, z, 0.0585814891280221);
  p = lsh_fma64(p, z, -0.06664511447381948);
  p = lsh_fma64(p, z, 0.07692183190826087);
  p = lsh_fma64(p, z, -0.09090904578123903);
  p = lsh_fma64(p, z, 0.11111111015256361);
  p = lsh_fma64(p, z, -0.14285714284666542);
  p = lsh_fma64(p, z, 0.1999999999999552);
  p = lsh_fma64(p, z, -0.3333333333333333);
  double r = lsh_fma64(q * z, p, q) + (upper ? kPi4 : 0.0);  // atan(num / den)
  r = swap ? kPi2 - r : r;
  r = x < 0.0 ? kPiHi - r : r;
  return y < 0.0 ? -r : r;
}

}  // namespace lsh
