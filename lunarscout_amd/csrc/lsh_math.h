// lsh_math.h -- float64 elementary functions of the lighting path, written for
// its inputs (finite, moderate magnitudes), shared by the HIP kernels and the
// host builds (no HIP headers: the tests compile it with a host compiler).
#pragma once

#if defined(__HIPCC__)
#define LSH_MHD __host__ __device__ inline
#else
#define LSH_MHD inline
#endif

namespace lsh {

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
  const double q = (upper ? num - den : num) / (upper ? num + den : den);
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
