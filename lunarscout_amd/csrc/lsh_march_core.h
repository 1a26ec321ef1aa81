// lsh_march_core.h -- the per-ray horizon march shared by the HIP kernels.
//
// Restates hierarchy_maximum (cuda_backend.py:656-874) and the per-pixel
// subpatch interpolation (cuda_backend.py:1040-1123, 965-987) as __host__
// __device__ code. The same source also compiles with a host C++ compiler (a
// CPU build of the kernel arithmetic used by the tests to check the GPU
// bit for bit); the product library only compiles it for the GPU.
//
// The step schedule is the reference's; the precision is split in two:
//   * the schedule state -- ray distance s, pixel position, tangent, cell-exit
//     distances and step lengths -- is evaluated in precision S (float64 in the
//     product kernel). Samples land wherever the schedule puts them, and a
//     sample that lands within float32 rounding (~1e-7 px) of a cell edge, or a
//     step that drifts by float32 rounding, sends a float32 schedule down another
//     branch than the float64 method: on rough terrain the sampling phase then
//     shifts by up to half a pixel and the horizon by up to several degrees;
//   * heights, bounds and slopes are evaluated in precision V (float32 in the
//     product kernel), arranged so rounding stays small where it matters:
//     pixel coordinates are carried relative to an integer base near the ray
//     start; heights enter the slope as differences to the observer, h - oz,
//     taken on the four DEM taps before the bilinear blend; and the level-0
//     slope uses (h - oz)(2R + h + oz), the form the reference already uses for
//     its cull bound (cuda_backend.py:777-795), instead of rounding R + h to
//     float32 (0.125 m) first. With S = float32 the ray distance is kept as an
//     unevaluated float32 pair (s_hi + s_lo).
// Each rearrangement is algebraically exact (Horner forms and fused
// multiply-adds round fewer times than the reference's separate operations);
// decisions, constants and the order in which samples and bound tests are taken
// are unchanged.

#pragma once

#include <stdint.h>

#include <type_traits>

#include "lsh_math.h"

#if defined(__HIPCC__)
#define LSH_HD __host__ __device__ inline
#define LSH_HD_MEMBER __host__ __device__
#else
#define LSH_HD inline
#define LSH_HD_MEMBER
#endif

namespace lsh {

constexpr float kBeamWidth = 0.004363323129985824f;  // 2*pi/1440 (cuda_backend.py:699)

struct MarchView {
  // level 0, the mips, then the 2x2 bound of every level (same layout), one allocation
  const float* base;
  uint32_t buffer_bytes;      // its size (the device loads are bounded by it)
  int32_t bound_offset;       // element offset of the bound copy from base
  const int32_t* level_offset;  // element offset of each level from base
  const int32_t* level_width;
  const int32_t* level_height;
  int32_t width, height, level_count;
  float map_resolution;  // (column + row resolution) / 2
  float radius;
  int32_t all_valid;     // 1 when every level-0 elevation is valid (no nodata tests needed)
  // map_resolution widened, and its binary exponent (res = f 2^e, f in
  // [0.5, 1)): uniform values the loop reads from scalar registers
  double resolution;
  int32_t resolution_exponent;
  double resolution_mantissa;  // res = resolution_mantissa 2^resolution_exponent
  // the level-2 bound copy (element offset from base, row length in cells) for
  // the guarded skips of lsh_march_pass; box_width 0 turns them off
  uint32_t box_offset;
  int32_t box_width;
};

// One pass's interpolated segment with its start relative to (bx, by), in the
// schedule precision S.
template <class S>
struct PassSegment {
  int32_t bx, by;
  S sx, sy;
  S a1, a2, a3, a4, b1, b2, b3, b4;
  S s_start, s_end, s_chord, c1, c2, c3;
};

struct MarchCounters {
  uint32_t samples, tests, steps;
};

// Step recorder of the march; the kernels use NoTrace, which compiles away.
// A row holds the reference's trace fields (hierarchy.py:29-52): s_km, true_m,
// level, cell_x, cell_y, pixel_x, pixel_y, bound, sample h - oz, slope,
// advance_km, action (0 descend, 1 cull, 2 nodata, 4 sample).
struct NoTrace {
  LSH_HD_MEMBER void row(float, float, int, int, int, float, float, float, float, float, float,
                         int) const {}
  // Called at every level-0 sample of a pass before it updates the running
  // maximum: sample index, the sample's cell (ix, iy) and clamped bilinear
  // fractions, h - oz, the true distance (m) and the near-field flag. A tracer
  // returning true ends the pass there (the adjoint's re-march to its winner).
  LSH_HD_MEMBER bool sample(int, int, int, double, double, double, double, bool) const {
    return false;
  }
};

LSH_HD bool lsh_valid(float v) {
  return !__builtin_isnan(v) && !__builtin_isinf(v) && v > -20000.0f;  // cuda_backend.py:43-45
}

template <class S>
LSH_HD S lsh_fma(S a, S b, S c) {
  return sizeof(S) == 4 ? static_cast<S>(__builtin_fmaf(float(a), float(b), float(c)))
                        : static_cast<S>(__builtin_fma(double(a), double(b), double(c)));
}

// evaluate and tangent (cuda_backend.py:62-84) at the start offset d, in one
// Horner pass with fused multiply-adds: the partial sums b3 = a4 d + a3,
// b2 = b3 d + a2, b1 = b2 d + a1 give the position b1 d + x0, and the same
// recurrence on them gives the derivative 4 a4 d^3 + 3 a3 d^2 + 2 a2 d + a1
// (c3 = a4 d + b3, c2 = c3 d + b2, c2 d + b1), with no per-ray 2 a2, 3 a3, 4 a4.
template <class S>
LSH_HD void lsh_quartic_and_tangent(S x0, S a1, S a2, S a3, S a4, S d, S& position,
                                    S& tangent) {
  const S b3 = lsh_fma(d, a4, a3);
  const S b2 = lsh_fma(d, b3, a2);
  const S b1 = lsh_fma(d, b2, a1);
  position = lsh_fma(d, b1, x0);
  const S c3 = lsh_fma(d, a4, b3);
  const S c2 = lsh_fma(d, c3, b2);
  tangent = lsh_fma(d, c2, b1);
}

template <class T>
LSH_HD T lsh_sqrt(T x) {
  return sizeof(T) == 4 ? static_cast<T>(__builtin_sqrtf(static_cast<float>(x)))
                        : static_cast<T>(__builtin_sqrt(static_cast<double>(x)));
}
template <class T>
LSH_HD T lsh_floor(T x) {
  return sizeof(T) == 4 ? static_cast<T>(__builtin_floorf(static_cast<float>(x)))
                        : static_cast<T>(__builtin_floor(static_cast<double>(x)));
}
// x 2^e, exact for the powers of four the schedule scales the resolution by
template <class T>
LSH_HD T lsh_ldexp(T x, int e) {
  return sizeof(T) == 4 ? static_cast<T>(__builtin_ldexpf(static_cast<float>(x), e))
                        : static_cast<T>(__builtin_ldexp(static_cast<double>(x), e));
}
// x = f 2^e with f in [0.5, 1): returns f, sets e
template <class T>
LSH_HD T lsh_frexp(T x, int& e) {
  e = 0;
  return sizeof(T) == 4 ? static_cast<T>(__builtin_frexpf(static_cast<float>(x), &e))
                        : static_cast<T>(__builtin_frexp(static_cast<double>(x), &e));
}
// the binary exponent e of x = f 2^e, f in [0.5, 1)
template <class T>
LSH_HD int lsh_exponent(T x) {
  int e = 0;
  if (sizeof(T) == 4)
    (void)__builtin_frexpf(static_cast<float>(x), &e);
  else
    (void)__builtin_frexp(static_cast<double>(x), &e);
  return e;
}
// max / min of numbers that are never NaN here (one v_max / v_min instruction)
template <class T>
LSH_HD T lsh_max(T a, T b) {
  return sizeof(T) == 4 ? static_cast<T>(__builtin_fmaxf(static_cast<float>(a), static_cast<float>(b)))
                        : static_cast<T>(__builtin_fmax(static_cast<double>(a), static_cast<double>(b)));
}
template <class T>
LSH_HD T lsh_min(T a, T b) {
  return sizeof(T) == 4 ? static_cast<T>(__builtin_fminf(static_cast<float>(a), static_cast<float>(b)))
                        : static_cast<T>(__builtin_fmin(static_cast<double>(a), static_cast<double>(b)));
}
// |a| (a source modifier on the device: no instruction of its own); the cell
// exit below never divides by a zero it would sign
template <class T>
LSH_HD T lsh_abs(T a) {
  return sizeof(T) == 4 ? static_cast<T>(__builtin_fabsf(static_cast<float>(a)))
                        : static_cast<T>(__builtin_fabs(static_cast<double>(a)));
}


// Elevation slope, on a sphere of radius R, of a point hd metres above an observer
// at height oz and `distance` metres away (cuda_backend.py:777-795, 812-830).
template <class V>
LSH_HD V lsh_sphere_slope(V hd, V oz, V distance, V radius, V two_observer_radius) {
  V sq = distance * distance;
  V rs = (V(2) * radius + V(2) * oz) + hd;  // 2R + h + oz with h = oz + hd
  V lz = (hd * rs - sq) / two_observer_radius;
  V lxx = sq - lz * lz;
  V lx = lxx > V(0) ? lsh_sqrt(lxx) : V(1e-6);
  return lz / lx;
}

// possible <= current for the cull test (cuda_backend.py:767-802) without
// forming the quotient: the bound's slope is lz / lx with lx = sqrt(lxx),
// lxx = sq - lz^2 (or lx = 1e-6 when lxx <= 0). For lxx > 0,
// lz / sqrt(lxx) <= current is decided on signed squares:
// lz |lz| <= current |current| lxx, which holds the cases of the comparison
// (lz <= 0 with current >= 0 always culls; lz > 0 with current <= 0 never
// does; otherwise lz^2 against current^2 lxx, the inequality reversed when both
// are negative). Exact arithmetic gives the same answer; rounding can differ
// only at ties. A wavefront in which no lane is near the observer and no lane
// has lxx <= 0 (the rule far out) takes that one comparison; otherwise every
// lane evaluates the near and the lx = 1e-6 cases as well.
template <class V>
LSH_HD bool lsh_bound_cannot_rise(V hd, V distance, bool near, V two_radius_oz,
                                  V inv_two_observer_radius, V current) {
  const V sq = distance * distance;
  const V rs = two_radius_oz + hd;  // 2R + h + oz with h = oz + hd
  const V lz = lsh_fma(hd, rs, -sq) * inv_two_observer_radius;
  const V lxx = lsh_fma(-lz, lz, sq);
  const bool degenerate = !(lxx > V(0));
  const bool ordinary = lz * lsh_abs(lz) <= (current * lsh_abs(current)) * lxx;
  if (!lsh_any_lane(near || degenerate)) return ordinary;
  lsh_keep_branch();
  const bool far_result = degenerate ? lz * V(1e6) <= current : ordinary;
  return near ? hd <= current * distance : far_result;  // near: distance >= 1 m
}

// Guarded skips. The cull test of a level-l cell (below) decides on the bound B
// of the cell's 2x2 block at the distance d = max(true_m - side, 1), side =
// 4^l res. Its slope is tan(theta) with sin(theta) = lz / d =
// (hd k - d^2) / (a d), k = 2R + h + oz, a = 2(R + oz), hd = B - oz:
// increasing in hd, and over d in [d_lo, d_hi] at most (hd k - d_lo d') / (a d')
// with d' = d_lo for hd >= 0 and d_hi for hd < 0 (each term bounded on its own;
// the slope is not monotone in d when hd < 0). A window of 16 x 16 level-l cells
// starting at (wx, wy) lies with the 2x2 blocks of its cells inside the
// 32 x 32 cells [16 (wx >> 4), +32) x [16 (wy >> 4), +32) whose maximum is the
// level-(l+2) bound at (wx >> 4, wy >> 4) (the mips are maxima). So when that
// bound, over the distances true_m -+ 32 side (more than any cell of the window
// lies from the position), cannot reach current - 1e-9 (a margin far above the
// rounding of either test), every position of the lane whose level-l cell lies
// in the window is culled (or nodata) by the method's own test; and when that
// distance range keeps the footprint in [4^l res, 4^(l+1) res), each such
// position starts at level l. lsh_guard_holds is that test for one lane.
template <class S, class V>
LSH_HD bool lsh_guard_holds(float box_bound, V oz, S true_m, S margin, V two_radius_oz,
                            V two_observer_radius, V current, S footprint_lo, S footprint_hi) {
  const V hd = static_cast<V>(box_bound) - oz;
  const S d_hi = true_m + margin;
  const S d_lo = lsh_max(true_m - margin, S(1));
  const S d_sel = hd >= V(0) ? d_lo : d_hi;
  const V num = lsh_fma(hd, two_radius_oz + hd, -(d_lo * d_sel));  // a d' sin(theta) bound
  const V ad = two_observer_radius * d_sel;
  const V c = current - V(1e-9);
  return (d_lo * static_cast<S>(kBeamWidth) >= footprint_lo) &
         (d_hi * static_cast<S>(kBeamWidth) < footprint_hi) & (lsh_abs(num) < ad) &
         (num * lsh_abs(num) <= (c * lsh_abs(c)) * lsh_fma(-num, num, ad * ad));
}

// a * b for 0 <= a, b < 2^24 (row index times row length: one full-rate
// 24-bit multiply on the device instead of a quarter-rate 32-bit one)
LSH_HD uint32_t lsh_mul24(int32_t a, int32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul24(static_cast<unsigned>(a), static_cast<unsigned>(b));
#else
  return static_cast<uint32_t>(a) * static_cast<uint32_t>(b);
#endif
}

// One element of the DEM's allocation by its offset from the base. On the GPU a
// buffer load with a 32-bit offset from a descriptor over the allocation
// (an out-of-range offset would read 0, never fault); the b32 load returns the
// raw 32 bits.
LSH_HD float lsh_load(const MarchView& d, uint32_t offset) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(d.base), static_cast<short>(0), static_cast<int>(d.buffer_bytes),
      0x00020000);
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                       rsrc, static_cast<int>(offset * 4u), 0, 0));
#else
  return d.base[offset];
#endif
}

// Ray distance s in km. Single precision keeps an unevaluated pair (hi + lo)
// updated by TwoSum; double precision keeps one value.
template <class S>
struct RayDistance {
  S hi, lo;
  LSH_HD void start(S v) { hi = v; lo = S(0); }
  LSH_HD void add(S a) {
    if (sizeof(S) == 8) {
      hi = hi + a;
      return;
    }
    S t = hi + a;
    S bp = t - hi;
    S e = (hi - (t - bp)) + (a - bp);
    lo = lo + e;
    hi = t + lo;
    lo = lo - (hi - t);
  }
  // (in double precision lo stays 0 and the forms reduce to plain hi)
  LSH_HD bool at_most(S v) const {
    return sizeof(S) == 8 ? hi <= v : (hi < v || (hi == v && lo <= S(0)));
  }
  LSH_HD bool below(S v) const {
    return sizeof(S) == 8 ? hi < v : (hi < v || (hi == v && lo < S(0)));
  }
  LSH_HD S minus(S v) const { return sizeof(S) == 8 ? hi - v : (hi - v) + lo; }
  LSH_HD S value() const { return sizeof(S) == 8 ? hi : hi + lo; }
  LSH_HD S times_1000() const {
    return sizeof(S) == 8 ? hi * S(1000) : hi * S(1000) + lo * S(1000);
  }
};

// Distance along the local tangent to the exit of the current cell, in km
// (cuda_backend.py:733-759): min over the axes of (far edge - p) / d, where an
// axis with |d| <= 1e-8 uses the inverse 1e30. The cell spans [x_lo, x_lo +
// side) x [y_lo, y_lo + side) in the segment's integer-based coordinates, where
// the position (px, py) lies; x_lo - px is then exact (both are multiples of
// px's unit in the last place and |x_lo - px| < side), so the far-edge offset
// is rounded once whichever way it is formed. The axes are compared by
// cross-multiplication, then the winner takes one division. A wavefront in
// which no lane has a tiny tangent (the rule) takes the edge the tangent
// points at as |edge - p| / |d|; otherwise every lane takes the general form,
// which gives the same quotient for the lanes without a tiny tangent.
// The common form: |edge - p| / |d| toward the edge the tangent points at, exact
// for every lane whose tangent is not tiny on either axis.
template <class S>
LSH_HD S lsh_cell_exit_common(int x_lo, int y_lo, int side, S px, S py, S dx, S dy) {
  const S adx = lsh_abs(dx), ady = lsh_abs(dy);
  const S nx = lsh_abs(static_cast<S>(dx > S(0) ? x_lo + side : x_lo) - px);
  const S ny = lsh_abs(static_cast<S>(dy > S(0) ? y_lo + side : y_lo) - py);
  const bool use_x = nx * ady <= ny * adx;
  return lsh_quotient(use_x ? nx : ny, use_x ? adx : ady);
}
// The general form, for any tangent; its tests are formed from opaque copies of
// the tangent so that they are not merged with (and do not add to) the tests of
// the common path.
template <class S>
LSH_HD S lsh_cell_exit_general(int x_lo, int y_lo, int side, S px, S py, S dx, S dy) {
  S ox = dx, oy = dy;
  lsh_keep_value(ox);
  lsh_keep_value(oy);
  const S gx = lsh_abs(ox), gy = lsh_abs(oy);
  const bool tx = !(gx > S(1e-8)), ty = !(gy > S(1e-8));
  // (edge - p) for the edge the tangent points at (the far one for a tiny tangent)
  const S ax = tx || ox > S(0) ? static_cast<S>(x_lo + side) - px : px - static_cast<S>(x_lo);
  const S ay = ty || oy > S(0) ? static_cast<S>(y_lo + side) - py : py - static_cast<S>(y_lo);
  // an axis with a tiny tangent never wins unless both are tiny
  const bool use_x = ty ? (tx ? ax <= ay : true) : (tx ? false : ax * gy <= ay * gx);
  const S num = use_x ? ax : ay;
  const S den = tx && ty ? S(1e-30) : (use_x ? gx : gy);
  return num / den;
}
// Either form: the common one in a wavefront where no lane has a tiny tangent.
template <class S>
LSH_HD S lsh_cell_exit(int x_lo, int y_lo, int side, S px, S py, S dx, S dy) {
  const bool tiny = !(lsh_abs(dx) > S(1e-8)) || !(lsh_abs(dy) > S(1e-8));
  if (!lsh_any_lane(tiny)) return lsh_cell_exit_common(x_lo, y_lo, side, px, py, dx, dy);
  return lsh_cell_exit_general(x_lo, y_lo, side, px, py, dx, dy);
}

// hierarchy_maximum (cuda_backend.py:656-874) for one DEM pass. S is the
// precision of the schedule state (ray distance, position, tangent, distances,
// steps); V the precision of heights and slopes. `current` is the running
// maximum slope (-1e30 when none yet); `best` receives
// (pass << 24) | level-0 sample index when a sample raises it.
//
// One outer iteration is one position of the schedule: evaluate the path, pick
// the start level, test the 2x2 bound there (descending only while the bound
// neither culls nor is at level 0, which is rare), then advance s by the cell
// skip or by the level-0 step. Both the skip and the step are formed for every
// lane and selected, so a wavefront executes one straight path per position.
// Constant divisions of the reference (x / 1000) are multiplications by 0.001.
template <class S, class V, bool kCap, bool kCount, class Tracer = NoTrace>
LSH_HD void lsh_march_pass(const PassSegment<S>& g, const MarchView& d, V oz, int pass,
                           V& current, int32_t& best, MarchCounters& cnt,
                           const Tracer& trace = Tracer()) {
  const S res = static_cast<S>(d.resolution);
  const V radius = static_cast<V>(d.radius);
  const V two_observer_radius = V(2) * (radius + oz);
  const V inv_two_observer_radius = V(1) / two_observer_radius;
  const V two_radius_oz = V(2) * radius + V(2) * oz;
  const S chord0 = g.s_chord * S(1000);
  const S min_step = (S(0.5) * res) * S(0.001);
  // half the level-0 cell side in km; the level-l fallback is this 4^l (exact)
  const S fallback0 = (res * S(0.5)) * S(0.001);
  const S res_mantissa = static_cast<S>(d.resolution_mantissa);
  const S res4 = res * S(4);  // the level-1 cell side (exact)
  const S far_step = (S(0.8) * res) * S(0.001);
  // the reference's px in [0, W-1) is floor(px) in [0, W-2]
  const S max_x = static_cast<S>(d.width - 2), max_y = static_cast<S>(d.height - 2);
  // the bilinear clamp at width - 1.0001 (cuda_backend.py:113-114), as a fraction
  const S clamp_tx = (static_cast<S>(d.width) - S(1.0001)) - static_cast<S>(d.width - 2);
  const S clamp_ty = (static_cast<S>(d.height) - S(1.0001)) - static_cast<S>(d.height - 2);
  const bool all_valid = d.all_valid != 0;
  const int top_level = d.level_count - 1;
  // Guarded skips (lsh_guard_holds): after a position at which every lane of
  // the wavefront starts at the same level and holds the guard, the next
  // kGuardSteps positions of every lane whose cell stays in its window are culls
  // at that level by the method's test, so they take only the path, the cell
  // exit and the skip; the window runs from one cell behind the lane's cell to
  // 14 ahead in the direction the lane moves. A wave-uniform count; lanes need
  // not reach the same cells. Not with a tracer (its rows name the bound and the
  // cull or nodata action) nor in the float32 builds (the 1e-9 margin is below
  // their rounding).
  constexpr bool kGuard =
      std::is_same<Tracer, NoTrace>::value && sizeof(S) == 8 && sizeof(V) == 8;
  constexpr int kGuardWindow = 16, kGuardSteps = 14;
  // a lane's window starts one cell behind its cell in the direction it moves
  auto window_start = [](int c, S tangent) {
    return tangent > S(0) ? c - 1 : c - (kGuardWindow - 2);
  };
  // The guard's distance range takes a position of the window (within
  // 15 sqrt(2) cells of the lane's position) to be at most 1.4 x that far in
  // true distance from the lane's: the chord polynomial's slope
  // |c1 + 2 c2 p + 3 c3 p^2| stays below 1.4 for planar distances p up to twice
  // the pass's path length (1.0 within a fraction of a percent in the polar
  // stereographic frame); a lane whose fit does not is never guarded.
  const S chord_reach = g.s_end * S(2000);
  const bool chord_tame = lsh_abs(g.c1) + lsh_abs(S(2) * g.c2) * chord_reach +
                              lsh_abs(S(3) * g.c3) * (chord_reach * chord_reach) <=
                          S(1.4);
  const bool guard_on = kGuard && d.box_width > 0;
  float box_bound = 0.0f;
  // the true distance in metres of a position away from the observer
  // (cuda_backend.py:690-698): the planar distance from the segment start and
  // chord0 + c1 p + c2 p^2 + c3 p^3 in Horner form
  auto far_true_m = [&](S px, S py) {
    const S plx = (px - g.sx) * res, ply = (py - g.sy) * res;
    const S planar = lsh_sqrt_any(lsh_fma(plx, plx, ply * ply));
    const S chord_h = lsh_fma(planar, lsh_fma(planar, g.c3, g.c2), g.c1);
    return lsh_fma(planar, chord_h, chord0);
  };
  RayDistance<S> s;
  s.start(lsh_max(g.s_start, S(0.001)));
  int sample_index = 0;
  // after a guard fails, the next kGuardBackoff positions that could test one do
  // not (a failing region tends to keep failing)
  constexpr int kGuardBackoff = 8;
  int cooldown = 0;
  while (s.at_most(g.s_end)) {
    // -- position (cuda_backend.py:672-705)
    const S delta = s.minus(g.s_start);
    S px, py, dx, dy;
    lsh_quartic_and_tangent(g.sx, g.a1, g.a2, g.a3, g.a4, delta, px, dx);
    lsh_quartic_and_tangent(g.sy, g.b1, g.b2, g.b3, g.b4, delta, py, dy);
    // the reference's px in [0, W-1), py in [0, H-1), tested on the absolute cell
    // index floor(px) + bx (exact: integers); NaN fails every comparison
    const S fx = lsh_floor(px), fy = lsh_floor(py);
    const S ax = fx + static_cast<S>(g.bx), ay = fy + static_cast<S>(g.by);
    if (!(ax >= S(0) && ay >= S(0) && ax <= max_x && ay <= max_y)) break;
    const int ix = static_cast<int>(ax), iy = static_cast<int>(ay);
    const bool near = s.below(S(0.5));
    // the far-field true distance; a lane near the observer (s < 0.5 km) uses s
    // itself, formed in the wavefronts that hold one (below)
    S true_m = far_true_m(px, py);
    // start level: the highest l in 1..top whose cell side 4^l res fits the beam
    // footprint, else 0 -- 0 exactly when footprint < 4 res (top >= 1 for any
    // DEM of at least 2 x 2). Otherwise, with footprint = f 2^ef and
    // res = r 2^er (f, r in [0.5, 1)), floor(log2(footprint / res)) =
    // ef - er - (f < r), exactly, and l is its floor half, clamped to top. No
    // loop, no rounding; formed only in a wavefront where a lane needs it.
    S footprint = true_m * static_cast<S>(kBeamWidth);
    int level = 0;
    auto start_level = [&]() {
      int footprint_exponent;
      const S footprint_mantissa = lsh_frexp(footprint, footprint_exponent);
      const int l = (footprint_exponent - d.resolution_exponent -
                     (footprint_mantissa < res_mantissa ? 1 : 0)) >> 1;
      return l < 0 ? 0 : (l < top_level ? l : top_level);
    };

    // -- hierarchy tests at this position (cuda_backend.py:706-868)
    // hierarchy_bilinear_bound (cuda_backend.py:47-60) -- the maximum of the
    // valid elevations of cells (cx..cx+1, cy..cy+1) clipped to the level,
    // -32000 when none is valid -- is read from the bound copy of the level,
    // where it is stored per cell (one load instead of four). The start
    // level's bound is loaded first, and the cell exit and fallback of that
    // level, which do not depend on it, are formed while the load is in flight.
    // A descent (the bound neither culls nor is at level 0, rare) tests the
    // finer levels in turn and forms the exit of the level it stops at. The
    // reference also ends a ray whose cell lies outside the level; the level
    // sizes are ceil(W / 4^l) x ceil(H / 4^l), so the cell of a position in
    // [0, W-2] x [0, H-2] lies inside every level and that test never fires.
    // The bound test at the start level, the cell exit and the cull test. The
    // start level is 0 for every lane of a wavefront out to the distance where
    // the footprint reaches 4 res (about 55 km at 60 m/px, the rule on these
    // DEMs): a wavefront with no lane there and none near the observer takes a
    // copy of the code specialised to level 0 and the far field -- the cell is
    // the level-0 cell, its bound sits at row iy, column ix of the level-0
    // bound copy (no level table), its side is 1 px, its fallback fallback0 --
    // and others the general one.
    int shift = 0, cx = ix, cy = iy;
    float maxh = 0.0f;
    S fallback = fallback0, exit_distance = S(0);
    bool nodata = false, culled = false;
    auto test_start_level = [&](auto at_level0) {
      constexpr bool kLevel0 = decltype(at_level0)::value;
      shift = kLevel0 ? 0 : level * 2;
      cx = kLevel0 ? ix : ix >> shift;
      cy = kLevel0 ? iy : iy >> shift;
      const int side = kLevel0 ? 1 : 1 << shift;
      const uint32_t bound_index =
          kLevel0 ? static_cast<uint32_t>(d.bound_offset) + lsh_mul24(iy, d.width) +
                        static_cast<uint32_t>(ix)
                  : static_cast<uint32_t>(d.bound_offset + d.level_offset[level] +
                                          lsh_mul24(cy, d.level_width[level]) + cx);
      maxh = lsh_load(d, bound_index);
      if (kGuard) {
        // the level-(l+2) bound a guard at this position tests, read with the
        // cell's (unconditionally, so that the cull test waits for its own load
        // only; a level without it reads an unused in-range value)
        const int wx = window_start(cx, dx), wy = window_start(cy, dy);
        const int gx = (wx > 0 ? wx : 0) >> 4, gy = (wy > 0 ? wy : 0) >> 4;
        if (kLevel0) {
          box_bound = lsh_load(d, d.box_offset + lsh_mul24(gy, d.box_width) +
                                      static_cast<uint32_t>(gx));
        } else {
          const int box_level = level + 2 <= top_level ? level + 2 : top_level;
          box_bound = lsh_load(d, static_cast<uint32_t>(d.bound_offset + d.level_offset[box_level]) +
                                      lsh_mul24(gy, d.level_width[box_level]) +
                                      static_cast<uint32_t>(gx));
        }
      }
      fallback = kLevel0 ? fallback0 : lsh_ldexp(fallback0, shift);
      const S cell_m = kLevel0 ? res : lsh_ldexp(res, shift);
      // the common cell exit (formed while the bound load is in flight) and
      // the cull test's one comparison; a wavefront in which a lane has a tiny
      // tangent, is near the observer or has lxx <= 0 takes the general forms
      // of both, for every lane
      const int x_lo = (kLevel0 ? ix : cx << shift) - g.bx;
      const int y_lo = (kLevel0 ? iy : cy << shift) - g.by;
      exit_distance = lsh_cell_exit_common(x_lo, y_lo, side, px, py, dx, dy);
      nodata = maxh < -20000.0f;
      const V hd0 = static_cast<V>(maxh) - oz;
      const V dist0 = static_cast<V>(lsh_max(true_m - cell_m, S(1)));
      const V sq0 = dist0 * dist0;
      const V lz0 = lsh_fma(hd0, two_radius_oz + hd0, -sq0) * inv_two_observer_radius;
      const V lxx0 = lsh_fma(-lz0, lz0, sq0);
      culled = lz0 * lsh_abs(lz0) <= (current * lsh_abs(current)) * lxx0;
      const bool tiny = !(lsh_abs(dx) > S(1e-8)) || !(lsh_abs(dy) > S(1e-8));
      // (no lane of a level-0 wavefront is near: the vote below sends those here)
      if (lsh_any_lane(tiny || (!kLevel0 && near) || !(lxx0 > V(0)))) {
        lsh_keep_branch();
        exit_distance = lsh_cell_exit_general(x_lo, y_lo, side, px, py, dx, dy);
        culled = lsh_bound_cannot_rise(hd0, dist0, near, two_radius_oz, inv_two_observer_radius,
                                       current);
      }
    };
    const bool level0_wave = !lsh_any_lane(near || footprint >= res4);
    if (level0_wave) {
      test_start_level(std::true_type{});
    } else {
      lsh_keep_branch();
      if (near) {
        true_m = s.times_1000();
        footprint = true_m * static_cast<S>(kBeamWidth);
      }
      level = start_level();
      test_start_level(std::false_type{});
    }
    const int start = level;  // (a descent below lowers level)
    if (kCount) {
      ++cnt.steps;
      if (nodata || culled || level != 0) ++cnt.tests;
    }
    if (!(nodata || culled || level == 0)) {
      do {
        trace.row(float(s.value()), float(true_m), level, cx, cy, float(px + g.bx),
                  float(py + g.by), maxh, 0.f, 0.f, 0.f, 0);
        --level;
        shift = level * 2;
        cx = ix >> shift;
        cy = iy >> shift;
        maxh = lsh_load(d, static_cast<uint32_t>(d.bound_offset + d.level_offset[level] +
                                                 lsh_mul24(cy, d.level_width[level]) + cx));
        nodata = maxh < -20000.0f;
        culled = lsh_bound_cannot_rise(
            static_cast<V>(maxh) - oz,
            static_cast<V>(lsh_max(true_m - lsh_ldexp(res, shift), S(1))), near, two_radius_oz,
            inv_two_observer_radius, current);
        if (kCount) {
          ++cnt.steps;
          if (nodata || culled || level != 0) ++cnt.tests;
        }
      } while (!(nodata || culled || level == 0));
      fallback = lsh_ldexp(fallback0, shift);
      exit_distance = lsh_cell_exit((cx << shift) - g.bx, (cy << shift) - g.by, 1 << shift, px,
                                    py, dx, dy);
    }
    const bool sampled = !nodata && !culled;  // at level 0
    // -- cell skip (cull / nodata) and level-0 step, selected per lane
    // the distance to the cell exit, or the fallback when the exit is not ahead
    // (cuda_backend.py:756-763): a cull or nodata skip moves 1e-4 km past it
    const S to_exit = exit_distance > S(0) ? exit_distance : fallback;
    const S skip = to_exit + S(0.0001);
    S advance = skip;
    if (sampled) {
      // bilinear (cuda_backend.py:107-132) of h - oz on the level-0 taps of the
      // bound just tested (inside the DEM: ix <= W - 2 and iy <= H - 2 here)
      const uint32_t r0 = lsh_mul24(iy, d.width) + static_cast<uint32_t>(ix);
      const uint32_t r1 = r0 + static_cast<uint32_t>(d.width);
      const float v00 = lsh_load(d, r0), v10 = lsh_load(d, r0 + 1);
      const float v01 = lsh_load(d, r1), v11 = lsh_load(d, r1 + 1);
      const bool taps_valid =
          all_valid || (lsh_valid(v00) && lsh_valid(v10) && lsh_valid(v01) && lsh_valid(v11));
      S sx_frac = px - fx, sy_frac = py - fy;
      if (ix == d.width - 2 && sx_frac > clamp_tx) sx_frac = clamp_tx;
      if (iy == d.height - 2 && sy_frac > clamp_ty) sy_frac = clamp_ty;
      const V tx = static_cast<V>(sx_frac), ty = static_cast<V>(sy_frac);
      const V d00 = V(v00) - oz, d10 = V(v10) - oz, d01 = V(v01) - oz, d11 = V(v11) - oz;
      const V top = d00 + tx * (d10 - d00);
      const V bottom = d01 + tx * (d11 - d01);
      const V hd = top + ty * (bottom - top);
      const V tm = static_cast<V>(true_m);
      // slope: (h - oz) / d near the observer, lz / lx on the sphere (cuda_backend.py:806-830)
      const V sq = tm * tm;
      const V lz = lsh_fma(hd, two_radius_oz + hd, -sq) * inv_two_observer_radius;
      const V lxx = lsh_fma(-lz, lz, sq);
      const V num = near ? hd : lz;
      const V den = near ? tm : (lxx > V(0) ? lsh_sqrt_any(lxx) : V(1e-6));
      V slope = lsh_ratio(num, den);
      if (!taps_valid) slope = V(-1e30);
      if (trace.sample(sample_index, ix, iy, static_cast<double>(tx), static_cast<double>(ty),
                       static_cast<double>(hd), static_cast<double>(tm), near))
        return;
      if (slope > current) {
        current = slope;
        best = (pass << 24) | sample_index;
      }
      if (kCount) ++cnt.samples;
      ++sample_index;
      // adaptive level-0 advance (cuda_backend.py:833-861)
      const S mag = lsh_sqrt_any(dx * dx + dy * dy);
      const S pixel_step = mag > S(1e-6) ? lsh_ratio(S(1), mag) : S(0.0005);
      const S margin = static_cast<S>(current - slope);
      const S margin_step = margin > S(0) ? ((margin * true_m) * S(1.732)) * S(0.001) : S(0);
      const S angular_step = (true_m * S(0.00151)) * S(0.001);
      S step = lsh_max(pixel_step, lsh_min(margin_step, angular_step));
      if (near) step = step * S(0.25);
      const S floor_step = (pass == 0 && true_m >= S(100)) ? far_step : min_step;
      step = lsh_max(step, floor_step);
      if (kCap) step = lsh_min(step, to_exit + S(0.000001));  // cuda_backend.py:636-640
      advance = step;
      trace.row(float(s.value()), float(true_m), 0, cx, cy, float(px + g.bx), float(py + g.by),
                maxh, float(hd), float(slope), float(advance), 4);
    } else {
      trace.row(float(s.value()), float(true_m), level, cx, cy, float(px + g.bx),
                float(py + g.by), maxh, 0.f, 0.f, float(skip), nodata ? 2 : 1);
    }
    s.add(advance);
    if constexpr (kGuard) {
      // a guard for the next positions when every lane starts at the same level
      // l (none near the observer) and the DEM has level l + 2 (current as this
      // position left it); its box must lie inside [0, W-2] x [0, H-2], where
      // the march never breaks
      const int gl = level0_wave ? 0 : lsh_uniform(start);
      const bool candidate =
          guard_on && (level0_wave || (gl + 2 <= top_level && !lsh_any_lane(start != gl || near)));
      if (candidate && cooldown > 0) {
        --cooldown;
      } else if (candidate) {
        const int gsh = 2 * gl;
        const int wx = window_start(ix >> gsh, dx), wy = window_start(iy >> gsh, dy);
        const S side_m = lsh_ldexp(res, gsh);
        const bool holds =
            chord_tame & (wx >= 0) & (wy >= 0) & (((wx + kGuardWindow) << gsh) <= d.width - 1) &
            (((wy + kGuardWindow) << gsh) <= d.height - 1) &
            lsh_guard_holds(box_bound, oz, true_m, S(2 * kGuardWindow) * side_m,
                            two_radius_oz, two_observer_radius, current,
                            gl > 0 ? side_m : S(0), side_m * S(4));
        if (lsh_any_lane(!holds)) {
          cooldown = kGuardBackoff;
        } else {
          // the guarded positions: the level-l cell of the position, its exit
          // and the skip. When a run ends with every lane still in its window,
          // the guard is tested again at the next position (which the new
          // window then covers too) instead of an ordinary position; a run
          // ends early (the position is then taken again as an ordinary one)
          // when a lane's cell leaves its window, and the guarded stretch ends
          // when a new guard does not hold
          const S fallback_l = lsh_ldexp(fallback0, gsh);
          const S margin = S(2 * kGuardWindow) * side_m;
          int gwx = wx, gwy = wy;
          for (;;) {
            bool left_window = false;
            for (int k = 0; k < kGuardSteps; ++k) {
              if (!s.at_most(g.s_end)) break;
              const S gdelta = s.minus(g.s_start);
              S gpx, gpy, gdx, gdy;
              lsh_quartic_and_tangent(g.sx, g.a1, g.a2, g.a3, g.a4, gdelta, gpx, gdx);
              lsh_quartic_and_tangent(g.sy, g.b1, g.b2, g.b3, g.b4, gdelta, gpy, gdy);
              const int kx = (static_cast<int>(lsh_floor(gpx)) + g.bx) >> gsh;
              const int ky = (static_cast<int>(lsh_floor(gpy)) + g.by) >> gsh;
              const bool inside = (static_cast<unsigned>(kx - gwx) < kGuardWindow) &
                                  (static_cast<unsigned>(ky - gwy) < kGuardWindow);
              if (lsh_any_lane(!inside)) {
                left_window = true;
                break;
              }
              const S gexit = lsh_cell_exit((kx << gsh) - g.bx, (ky << gsh) - g.by, 1 << gsh,
                                            gpx, gpy, gdx, gdy);
              if (kCount) {
                ++cnt.steps;
                ++cnt.tests;
              }
              s.add((gexit > S(0) ? gexit : fallback_l) + S(0.0001));
            }
            if (lsh_any_lane(left_window)) break;
            const bool active = s.at_most(g.s_end);
            if (!lsh_any_lane(active)) break;
            // the guard again, at the next position of the lanes still marching
            // (none is near the observer: s only grows)
            const S gdelta = s.minus(g.s_start);
            S gpx, gpy, gdx, gdy;
            lsh_quartic_and_tangent(g.sx, g.a1, g.a2, g.a3, g.a4, gdelta, gpx, gdx);
            lsh_quartic_and_tangent(g.sy, g.b1, g.b2, g.b3, g.b4, gdelta, gpy, gdy);
            const int kx = (static_cast<int>(lsh_floor(gpx)) + g.bx) >> gsh;
            const int ky = (static_cast<int>(lsh_floor(gpy)) + g.by) >> gsh;
            const int nwx = window_start(kx, gdx), nwy = window_start(ky, gdy);
            const S g_true_m = far_true_m(gpx, gpy);
            const int box_level = gl + 2;
            const float g_box = lsh_load(
                d, static_cast<uint32_t>(d.bound_offset + d.level_offset[box_level]) +
                       lsh_mul24((nwy > 0 ? nwy : 0) >> 4, d.level_width[box_level]) +
                       static_cast<uint32_t>((nwx > 0 ? nwx : 0) >> 4));
            const bool again =
                (!active) |
                (chord_tame & (nwx >= 0) & (nwy >= 0) &
                 (((nwx + kGuardWindow) << gsh) <= d.width - 1) &
                 (((nwy + kGuardWindow) << gsh) <= d.height - 1) &
                 lsh_guard_holds(g_box, oz, g_true_m, margin, two_radius_oz, two_observer_radius,
                                 current, gl > 0 ? side_m : S(0), side_m * S(4)));
            if (lsh_any_lane(!again)) break;
            if (active) {
              const S gexit = lsh_cell_exit((kx << gsh) - g.bx, (ky << gsh) - g.by, 1 << gsh, gpx,
                                            gpy, gdx, gdy);
              if (kCount) {
                ++cnt.steps;
                ++cnt.tests;
              }
              s.add((gexit > S(0) ? gexit : fallback_l) + S(0.0001));
            }
            gwx = nwx;
            gwy = nwy;
          }
        }
      }
    }
  }
}

LSH_HD int lsh_clamp_center(int requested, int dem_size, int sub) {
  int half = sub / 2;
  int maximum = dem_size - half;
  if (maximum < half) maximum = half;
  return requested < half ? half : (requested > maximum ? maximum : requested);
}

// bilinear (cuda_backend.py:107-132) of a DEM at an integer observer pixel, in
// precision V (the clamp at size - 1.0001 interpolates the last row/column).
template <class V>
LSH_HD V lsh_observer_terrain(const float* e, int w, int h, int column_i, int row_i) {
  V column = static_cast<V>(column_i), row = static_cast<V>(row_i);
  const V cmax = static_cast<V>(w) - V(1.0001);
  const V rmax = static_cast<V>(h) - V(1.0001);
  column = lsh_min(lsh_max(column, V(0)), cmax);
  row = lsh_min(lsh_max(row, V(0)), rmax);
  int x0 = static_cast<int>(lsh_floor(column)), y0 = static_cast<int>(lsh_floor(row));
  int x1 = x0 + 1 < w ? x0 + 1 : w - 1, y1 = y0 + 1 < h ? y0 + 1 : h - 1;
  float q00 = e[static_cast<int64_t>(y0) * w + x0], q10 = e[static_cast<int64_t>(y0) * w + x1];
  float q01 = e[static_cast<int64_t>(y1) * w + x0], q11 = e[static_cast<int64_t>(y1) * w + x1];
  if (!(lsh_valid(q00) && lsh_valid(q10) && lsh_valid(q01) && lsh_valid(q11))) return V(-32000);
  V tx = column - static_cast<V>(x0), ty = row - static_cast<V>(y0);
  V top = V(q00) + tx * (V(q10) - V(q00));
  V bottom = V(q01) + tx * (V(q11) - V(q01));
  return top + ty * (bottom - top);
}

// Per-pixel subpatch selection (cuda_backend.py:1040-1095).
struct PixelSelection {
  int32_t i00, i10, i01, i11;
  float tx, ty;
  int32_t cl, cr, ct, cb;  // clamped centers relative to the tile origin
};

LSH_HD PixelSelection lsh_select(int column, int row, int tile_column, int tile_row,
                                 int tile_width, int sub, int primary_w, int primary_h) {
  PixelSelection s;
  const int count = tile_width / sub + 2;
  float gx = static_cast<float>(column - sub / 2) / static_cast<float>(sub) + 1.0f;
  float gy = static_cast<float>(row - sub / 2) / static_cast<float>(sub) + 1.0f;
  int left = static_cast<int>(gx), top = static_cast<int>(gy);
  s.tx = gx - static_cast<float>(left);
  s.ty = gy - static_cast<float>(top);
  if (left < 0) { left = 0; s.tx = 0.0f; }
  if (top < 0) { top = 0; s.ty = 0.0f; }
  if (left > count - 2) { left = count - 2; s.tx = 1.0f; }
  if (top > count - 2) { top = count - 2; s.ty = 1.0f; }
  s.i00 = top * count + left;
  s.i10 = s.i00 + 1;
  s.i01 = (top + 1) * count + left;
  s.i11 = s.i01 + 1;
  const int rl = tile_column + (left - 1) * sub + sub / 2;
  const int rt = tile_row + (top - 1) * sub + sub / 2;
  s.cl = lsh_clamp_center(rl, primary_w, sub) - tile_column;
  s.cr = lsh_clamp_center(rl + sub, primary_w, sub) - tile_column;
  s.ct = lsh_clamp_center(rt, primary_h, sub) - tile_row;
  s.cb = lsh_clamp_center(rt + sub, primary_h, sub) - tile_row;
  return s;
}

// interpolate_subpatch_field (cuda_backend.py:965-987) of one pass, in precision S.
// p00..p11 point at the 18 fields of the four corner segments; start
// coordinates are taken relative to the integer base floor(p00[0]),
// floor(p00[1]) before the shifts and the bilinear blend (exact: the corners
// lie within a few pixels).
template <class S>
LSH_HD PassSegment<S> lsh_interpolate(const float* p00, const float* p10, const float* p01,
                                      const float* p11, const PixelSelection& sel, int column,
                                      int row, float ratio) {
  PassSegment<S> g;
  const S tx = sel.tx, ty = sel.ty;
  g.bx = static_cast<int32_t>(__builtin_floorf(p00[0]));
  g.by = static_cast<int32_t>(__builtin_floorf(p00[1]));
  const S fbx = static_cast<S>(g.bx), fby = static_cast<S>(g.by);
  const S r = static_cast<S>(ratio);
  const S sxl = static_cast<S>(column - sel.cl) * r;
  const S sxr = static_cast<S>(column - sel.cr) * r;
  const S syt = static_cast<S>(row - sel.ct) * r;
  const S syb = static_cast<S>(row - sel.cb) * r;
  auto fma = [](S a, S b, S c) {
    return sizeof(S) == 4 ? static_cast<S>(__builtin_fmaf(float(a), float(b), float(c)))
                          : static_cast<S>(__builtin_fma(double(a), double(b), double(c)));
  };
  auto blend = [&](S v00, S v10, S v01, S v11) {
    S upper = fma(v10 - v00, tx, v00);
    S lower = fma(v11 - v01, tx, v01);
    return fma(lower - upper, ty, upper);
  };
  g.sx = blend((S(p00[0]) - fbx) + sxl, (S(p10[0]) - fbx) + sxr, (S(p01[0]) - fbx) + sxl,
               (S(p11[0]) - fbx) + sxr);
  g.sy = blend((S(p00[1]) - fby) + syt, (S(p10[1]) - fby) + syt, (S(p01[1]) - fby) + syb,
               (S(p11[1]) - fby) + syb);
#define LSH_FIELD(f) blend(S(p00[f]), S(p10[f]), S(p01[f]), S(p11[f]))
  g.a1 = LSH_FIELD(4); g.a2 = LSH_FIELD(5); g.a3 = LSH_FIELD(6); g.a4 = LSH_FIELD(7);
  g.b1 = LSH_FIELD(8); g.b2 = LSH_FIELD(9); g.b3 = LSH_FIELD(10); g.b4 = LSH_FIELD(11);
  g.s_start = LSH_FIELD(12); g.s_end = LSH_FIELD(13); g.s_chord = LSH_FIELD(14);
  g.c1 = LSH_FIELD(15); g.c2 = LSH_FIELD(16); g.c3 = LSH_FIELD(17);
#undef LSH_FIELD
  return g;
}

}  // namespace lsh
