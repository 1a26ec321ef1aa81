// lsh_lighting_core.h -- per-pixel lighting from a stored horizon row, shared by
// the HIP kernel and the multi-threaded host path (the "cpu" backend).
//
// Restates (paths under src/lunarscout/_numba_horizon/):
//   * _pixel_frame and _azimuth_elevation_deg (psr.py:29-58) in float64;
//   * _sun_fraction_reference / BuilderSunFraction (lightmap.py:32-73,
//     lightmap_cuda.py:146-182) in float32, with the slice table of lightmap.py:15-29;
//   * the body-centre margin el - lerp(horizon) (lightmap_cpu.py:121-143,
//     lightmap.py:116-120) and the PSR upper-limb test (psr.py:94-140,
//     psr_cuda.py:171-186).
#pragma once

#include <cmath>
#include <cstdint>

#include "lsh_internal.h"
#include "lsh_math.h"

#if defined(__HIPCC__)
#define LSH_LHD __host__ __device__ inline
#else
#define LSH_LHD inline
#endif

namespace lsh {

constexpr double kPi = 3.14159265358979323846;
constexpr int kPatchPixels = 128 * 128;

struct Frame {
  double rot[3][3];  // rot[k][j]: component k of basis j (east, north, up)
  double tr[3];
};

// _pixel_frame (psr.py:29-47) in float64.
LSH_LHD Frame pixel_frame(const LightArgs& a, int row, int column) {
  const double* t = a.geo_transform;
  const double* p = a.projection;
  double x = t[0] + t[1] * column + t[2] * row;
  double y = t[3] + t[4] * column + t[5] * row;
  double xp = x - p[4], yp = y - p[5];
  double rho = hypot(xp, yp);
  double lat, lon;
  if (rho < 1e-9) {
    lat = p[1];
    lon = p[2];
  } else {
    double c = 2.0 * atan2(rho, 2.0 * p[3] * p[0]);
    double sc, cc, so, co;
    sincos(c, &sc, &cc);
    sincos(p[1], &so, &co);
    lat = asin(cc * so + yp * sc * co / rho);
    lon = p[2] + atan2(xp * sc, rho * co * cc - yp * so * sc);
  }
  double sl, cl, sn, cn;
  sincos(lat, &sl, &cl);
  sincos(lon, &sn, &cn);
  double up[3] = {cl * cn, cl * sn, sl};
  double east[3] = {-sn, cn, 0.0};
  double north[3] = {-sl * cn, -sl * sn, cl};
  double radius = p[0] + static_cast<double>(a.elevation[static_cast<int64_t>(row) * a.dem_width + column]);
  double obs[3] = {radius * up[0], radius * up[1], radius * up[2]};
  Frame f;
  for (int k = 0; k < 3; ++k) {
    f.rot[k][0] = east[k];
    f.rot[k][1] = north[k];
    f.rot[k][2] = up[k];
  }
  for (int j = 0; j < 3; ++j) {
    double acc = 0.0;
    for (int k = 0; k < 3; ++k) acc += obs[k] * f.rot[k][j];
    f.tr[j] = -acc;
  }
  return f;
}

// _azimuth_elevation_deg (psr.py:50-58) in float64: the body vector in the
// pixel's (east, north, up) frame and its horizontal length, then az = atan2(e, n)
// in degrees wrapped into [0, 360) as Python's % does (atan2 lies in
// (-180, 180], so one addition of 360 to a negative angle is that remainder),
// el = atan2(u, |(e, n)|).
LSH_LHD void body_vector(const Frame& f, const double* v, double* e, double& horizontal) {
  for (int j = 0; j < 3; ++j)
    e[j] = v[0] * f.rot[0][j] + v[1] * f.rot[1][j] + v[2] * f.rot[2][j] + f.tr[j];
  horizontal = lsh_sqrt_any(e[0] * e[0] + e[1] * e[1]);
}
LSH_LHD void azimuth_elevation_of(const double* e, double horizontal, double& az, double& el) {
  double a = atan2_finite(e[0], e[1]) * (180.0 / kPi);
  if (a < 0.0) a += 360.0;
  az = a;
  el = atan2_finite(e[2], horizontal) * (180.0 / kPi);
}
LSH_LHD void azimuth_elevation(const Frame& f, const double* v, double& az, double& el) {
  double e[3], horizontal;
  body_vector(f, v, e, horizontal);
  azimuth_elevation_of(e, horizontal, az, el);
}

// Decisions from horizon extremes. With lo and hi the lowest and highest of the
// bins a body's test can read (every one finite) and c the tallest slice
// half-height, a disk whose centre elevation is below lo - c - 2e-3 deg has
// every slice top below every interpolated horizon, so the walk adds nothing
// (0); one above hi + c + 2e-3 deg has every slice wholly above, so the walk
// adds every full slice (max_photons exactly, 1). The margin is far above the
// float32 rounding of the walk. Likewise a margin el - lerp(horizon) lies below
// a threshold when el < lo + threshold - 1e-4 deg and above it when
// el > hi + threshold + 1e-4 deg. Every test is taken on the body vector before
// any arctangent: el < t <=> u < |(e, n)| tan(t) for |t| < 90 deg. Bins that are
// not all finite, or a bound beyond +-89 deg, decide nothing.
struct RowExtremes {
  float lo, hi;  // lowest and highest bin
  bool finite;   // every bin finite
};
template <class Row>
LSH_LHD RowExtremes row_extremes(const Row& hz, int N) {
  RowExtremes x{1e30f, -1e30f, true};
  for (int b = 0; b < N; ++b) {
    const float v = hz[b];
    x.finite = x.finite && v > -1e30f && v < 1e30f;
    x.lo = v < x.lo ? v : x.lo;
    x.hi = v > x.hi ? v : x.hi;
  }
  return x;
}
// The disk walk also reads past the row's extremes: a disk centre within
// sun_half_angle + half a bin of azimuth 0 puts the walk's first position left
// of bin 0 (a negative fraction), and its first slices extrapolate from the
// pairs (bin 0, bin 1) or (bin N-1, bin 0) back to at most 2 h[0] - h[1] or
// 2 h[N-1] - h[0]. Those two values widen the extremes the disk tests use for
// the bins at azimuth 0.
template <class Row>
LSH_LHD void add_walk_extrapolation(const Row& hz, int N, RowExtremes& x) {
  const float h0 = hz[0], h1 = hz[1], hn = hz[N - 1];
  const float a = h0 + (h0 - h1), b = hn + (hn - h0);
  x.finite = x.finite && a > -1e30f && a < 1e30f && b > -1e30f && b < 1e30f;
  x.lo = a < x.lo ? a : x.lo;
  x.lo = b < x.lo ? b : x.lo;
  x.hi = a > x.hi ? a : x.hi;
  x.hi = b > x.hi ? b : x.hi;
}
// tan of an elevation bound in degrees, or the infinity that decides nothing
LSH_LHD double tan_bound(bool usable, double deg, double none) {
  return usable && deg > -89.0 && deg < 89.0 ? tan(deg * (kPi / 180.0)) : none;
}
// Sector decisions: the same tests on the extremes of the bins of one of 16
// azimuth sectors of 22.5 deg, widened by 4 bins (1 deg) on either side -- more
// than the disk walk reaches beyond the centre's azimuth (-0.4 .. +0.86 deg) or
// the horizon lerp (one bin), and than a body on a sector edge can be assigned
// to the neighbouring sector by rounding. The sector comes from the signs and
// ratio of the body's east and north components (no arctangent). The tangents
// are kept as float32 (the tables live in LDS); their rounding moves a bound
// by at most 2e-6 deg, inside the margins (2e-3 deg for the disk, 1e-4 deg
// for a margin's sign). Rows of another bin count than 1440 keep the
// whole-row decisions in every sector.
constexpr int kSectors = 16;
// A series decides only when it is long enough to repay reading the whole row
// first (the tables cost about one pass over the row's bins)
constexpr int kDecideMinTimes = 512;
struct SectorTable {
  float* dark;  // tan of the sector's "below" bound, element j at dark[j * stride]
  float* lit;
  int stride;
};
LSH_LHD int sector_of(double east, double north) {
  constexpr double kTan225 = 0.41421356237309504880, kTan675 = 2.41421356237309504880;
  const double ae = east < 0.0 ? -east : east, an = north < 0.0 ? -north : north;
  const int quadrant = east >= 0.0 ? (north >= 0.0 ? 0 : 1) : (north < 0.0 ? 2 : 3);
  int sub = (ae >= kTan225 * an) + (ae >= an) + (ae >= kTan675 * an);  // angle from the N-S axis
  if (quadrant == 1 || quadrant == 3) sub = 3 - sub;
  return 4 * quadrant + sub;
}
// below: deg offset of the "below" bound from a sector's lowest bin (negative),
// above: of the "above" bound from its highest bin
// (kS sectors of 360 / kS deg: 16, or 8 -- pairs of the 16, sector_of / 2)
template <int kS = kSectors, class Row>
LSH_LHD void fill_sectors(const Row& hz, int N, double below, double above, bool disk_walk,
                          SectorTable t) {
  if (N != 1440) {
    RowExtremes x = row_extremes(hz, N);
    if (disk_walk) add_walk_extrapolation(hz, N, x);
    const float d = static_cast<float>(tan_bound(x.finite, static_cast<double>(x.lo) + below, -INFINITY));
    const float l = static_cast<float>(tan_bound(x.finite, static_cast<double>(x.hi) + above, INFINITY));
    for (int j = 0; j < kS; ++j) {
      t.dark[j * t.stride] = d;
      t.lit[j * t.stride] = l;
    }
    return;
  }
  constexpr int kWidth = 1440 / kS, kExtra = 4;
  for (int j = 0; j < kS; ++j) {
    RowExtremes x{1e30f, -1e30f, true};
    for (int k = -kExtra; k < kWidth + kExtra; ++k) {
      int b = j * kWidth + k;
      b = b < 0 ? b + 1440 : (b >= 1440 ? b - 1440 : b);
      const float v = hz[b];
      x.finite = x.finite && v > -1e30f && v < 1e30f;
      x.lo = v < x.lo ? v : x.lo;
      x.hi = v > x.hi ? v : x.hi;
    }
    // the sectors on either side of azimuth 0 (a centre there may fall in either)
    if (disk_walk && (j == 0 || j == kS - 1)) add_walk_extrapolation(hz, N, x);
    t.dark[j * t.stride] =
        static_cast<float>(tan_bound(x.finite, static_cast<double>(x.lo) + below, -INFINITY));
    t.lit[j * t.stride] =
        static_cast<float>(tan_bound(x.finite, static_cast<double>(x.hi) + above, INFINITY));
  }
}

template <bool kAzimuthMajor>
struct HorizonRow {
  const float* base;
  int N;
  LSH_LHD float operator[](int bin) const {
    return kAzimuthMajor ? base[static_cast<int64_t>(bin) * kPatchPixels] : base[bin];
  }
};

// Interpolated horizon at an azimuth (psr.py:127-134), float64 position.
template <class Row>
LSH_LHD double horizon_at(const Row& hz, int N, double az) {
  double pos = az * (N / 360.0);
  double fl = floor(pos);
  // az in [0, 360] puts lower in [0, N]: the remainder (an integer division)
  // only when a lane of the wavefront is outside [0, N)
  int lower = static_cast<int>(fl);
  if (lsh_any_lane(lower < 0 || lower >= N)) {
    lower %= N;
    if (lower < 0) lower += N;
  }
  int upper = lower + 1 == N ? 0 : lower + 1;
  double f = pos - fl;
  double lo = hz[lower];
  return lo + f * (static_cast<double>(hz[upper]) - lo);
}

// The disk walk's scale for an N-bin horizon row: bins per degree, half a bin
// in degrees and the slice spacing in bins. The reference walks 1440 bins with
// the constants x4.0, 0.125 and 0.135 (lightmap.py:15-29, 39-43), which these
// are exactly at N = 1440; another N scales them to its bin width, so the walk
// never leaves the row (a fixed x4.0 would index up to 4 x 360 bins).
struct DiskWalk {
  float per_degree, half_bucket, step;
};
LSH_LHD DiskWalk disk_walk(int N) {
  const float b = static_cast<float>(N) / 360.0f;
  return DiskWalk{b, 0.5f / b, 0.135f * (b * 0.25f)};
}

// BuilderSunFraction (lightmap.py:32-73) in float32.
template <class Row>
LSH_LHD float sun_fraction(const Row& hz, int N, float az_deg, float el_deg,
                                     const LightArgs& a) {
  const DiskWalk w = disk_walk(N);
  float left_pos = (az_deg - a.sun_half_angle - w.half_bucket) * w.per_degree;
  int left = static_cast<int>(left_pos);  // truncation toward zero
  float frac = left_pos - static_cast<float>(left);
  if (left < 0) left += N;
  else if (left >= N) left -= N;
  int right = left + 1;
  if (right >= N) right = 0;
  float le = hz[left], re = hz[right];
  if (w.step <= 0.135f && frac >= 0.0f) {
    // With frac >= 0 the walk interpolates within bins left .. left + 4
    // (frac + 15 steps < 4 bins); a truncated negative position (az below half
    // the disk plus half a bin) extrapolates left of bin 0 instead and takes the
    // walk. A disk whose lowest slice top clears the highest of them by 1e-3
    // deg adds every full slice height (photons = max_photons exactly: doubling
    // commutes with rounding), one whose highest slice top stays 1e-3 deg under
    // the lowest adds nothing; rounding of the interpolation is ~1e-5 deg, so
    // the walk would give the same 1 or 0. Non-finite bins take the walk.
    int b = right;
    float lo = le < re ? le : re, hi = le < re ? re : le;
    for (int k = 0; k < 3; ++k) {
      b = b + 1 == N ? 0 : b + 1;
      const float v = hz[b];
      lo = v < lo ? v : lo;
      hi = v > hi ? v : hi;
    }
    const float c_max = a.half_circle[7] > a.half_circle[8] ? a.half_circle[7] : a.half_circle[8];
    const bool finite = lo > -1e30f && hi < 1e30f;
    if (finite && el_deg - c_max > hi + 1e-3f) return 1.0f;
    if (finite && el_deg + c_max + 1e-3f < lo) return 0.0f;
  }
  float delta = re - le;
  float photons = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const float col = a.half_circle[i];
    const float horizon = frac * delta + le;
    const float top = el_deg + col;
    if (horizon < top) {
      float ad = top - horizon;
      const float ch = col + col;
      if (ad > ch) ad = ch;
      photons = photons + ad;
    }
    frac = frac + w.step;
    if (frac >= 1.0f) {
      right += 1;
      if (right >= N) right = 0;
      le = re;
      re = hz[right];
      delta = re - le;
      frac = frac - 1.0f;
    }
  }
  return photons / a.max_photons;
}


// The three independent series of one pixel: the Sun walk (per-time fractions
// and their mean), the Earth margins (and the visible fraction) and the PSR
// limb test. Each writes only its own outputs, so they may run as separate
// work items.
constexpr unsigned kSunSeries = 1u, kEarthSeries = 2u, kPsrSeries = 4u;
constexpr unsigned kAllSeries = kSunSeries | kEarthSeries | kPsrSeries;

// The outputs of one pixel of one tile for the series in `series`: each time
// series walked in order.
template <bool kAzimuthMajor>
LSH_LHD void light_pixel(const LightArgs& a, int tile, int pixel, lsh_tile origin,
                         SectorTable sectors, unsigned series = kAllSeries) {
  const int line = pixel / 128, sample = pixel % 128;
  const int row = origin.tile_row + line, column = origin.tile_column + sample;
  const int64_t tile_pixels = static_cast<int64_t>(tile) * kPatchPixels;
  const bool inside = row < a.dem_height && column < a.dem_width;
  const bool sun = (series & kSunSeries) != 0, earth = (series & kEarthSeries) != 0;
  const bool limb = (series & kPsrSeries) != 0;
  if (!inside) {  // edge padding (psr_cuda.py:74-76)
    if (sun && a.sun_bytes)
      for (int t = 0; t < a.sun_count; ++t)
        a.sun_bytes[(static_cast<int64_t>(tile) * a.sun_count + t) * kPatchPixels + pixel] = 0;
    if (sun && a.sun_fraction)
      for (int t = 0; t < a.sun_count; ++t)
        a.sun_fraction[(static_cast<int64_t>(tile) * a.sun_count + t) * kPatchPixels + pixel] = 0.0f;
    if (earth && a.earth_margin)
      for (int t = 0; t < a.earth_count; ++t)
        a.earth_margin[(static_cast<int64_t>(tile) * a.earth_count + t) * kPatchPixels + pixel] =
            -INFINITY;
    if (sun && a.sun_mean) a.sun_mean[tile_pixels + pixel] = 0.0f;
    if (earth && a.earth_visible) a.earth_visible[tile_pixels + pixel] = 0.0f;
    if (limb && a.psr) a.psr[tile_pixels + pixel] = 0;
    if (limb && a.psr_margin) a.psr_margin[tile_pixels + pixel] = -INFINITY;
    return;
  }
  const int N = a.N;
  HorizonRow<kAzimuthMajor> hz;
  hz.N = N;
  hz.base = kAzimuthMajor ? a.horizons + static_cast<int64_t>(tile) * kPatchPixels * N + pixel
                          : a.horizons + (tile_pixels + pixel) * N;
  const Frame f = pixel_frame(a, row, column);

  if (sun && a.sun_count > 0 && (a.sun_bytes || a.sun_fraction || a.sun_mean)) {
    float acc = 0.0f;
    float c = 0.0f;  // the tallest slice half-height
    for (int i = 0; i < 16; ++i) c = a.half_circle[i] > c ? a.half_circle[i] : c;
    const bool decide = a.sun_count >= kDecideMinTimes;
    if (decide)
      fill_sectors(hz, N, -static_cast<double>(c) - 2e-3, static_cast<double>(c) + 2e-3, true,
                   sectors);
    for (int t = 0; t < a.sun_count; ++t) {
      double e[3], horizontal;
      body_vector(f, a.sun + 3 * t, e, horizontal);
      const int j = decide ? sector_of(e[0], e[1]) * sectors.stride : 0;
      float fr;
      if (decide && e[2] < horizontal * static_cast<double>(sectors.dark[j])) {
        fr = 0.0f;
      } else if (decide && e[2] > horizontal * static_cast<double>(sectors.lit[j])) {
        fr = 1.0f;
      } else {
        double az, el;
        azimuth_elevation_of(e, horizontal, az, el);
        fr = sun_fraction(hz, N, static_cast<float>(az), static_cast<float>(el), a);
      }
      acc += fr;
      int64_t o = (static_cast<int64_t>(tile) * a.sun_count + t) * kPatchPixels + pixel;
      if (a.sun_bytes) a.sun_bytes[o] = static_cast<uint8_t>(255.0f * fr);
      if (a.sun_fraction) a.sun_fraction[o] = fr;
    }
    if (a.sun_mean) a.sun_mean[tile_pixels + pixel] = acc / static_cast<float>(a.sun_count);
  }
  if (earth && a.earth_count > 0 && (a.earth_margin || a.earth_visible)) {
    int visible = 0;
    // the visible count alone is decided for the whole row where the Earth's
    // centre clears the row's extremes; the margins need every value
    const bool decide = a.earth_margin == nullptr && a.earth_count >= kDecideMinTimes;
    if (decide) fill_sectors(hz, N, -1e-4, 1e-4, false, sectors);
    for (int t = 0; t < a.earth_count; ++t) {
      double e[3], horizontal;
      body_vector(f, a.earth + 3 * t, e, horizontal);
      if (decide) {
        const int j = sector_of(e[0], e[1]) * sectors.stride;
        if (e[2] < horizontal * static_cast<double>(sectors.dark[j])) continue;
        if (e[2] > horizontal * static_cast<double>(sectors.lit[j])) {
          ++visible;
          continue;
        }
      }
      double az, el;
      azimuth_elevation_of(e, horizontal, az, el);
      double m = el - horizon_at(hz, N, az);
      visible += m >= 0.0;
      if (a.earth_margin)
        a.earth_margin[(static_cast<int64_t>(tile) * a.earth_count + t) * kPatchPixels + pixel] =
            static_cast<float>(m);
    }
    if (a.earth_visible)
      a.earth_visible[tile_pixels + pixel] = static_cast<float>(visible) / a.earth_count;
  }
  if (limb && a.psr_count > 0 && (a.psr || a.psr_margin)) {
    // the mask alone is decided by the first vector that lights the limb
    // (psr.py:136-139); the margin needs every vector
    const bool mask_only = a.psr_margin == nullptr;
    // with the mask alone, a vector below the row's lowest bin (less the
    // threshold) cannot light the limb and one above its highest does
    const bool decide = mask_only && a.psr_count >= kDecideMinTimes;
    if (decide) fill_sectors(hz, N, a.psr_threshold - 1e-4, a.psr_threshold + 1e-4, false, sectors);
    double best = -1e300;
    for (int t = 0; t < a.psr_count; ++t) {
      double e[3], horizontal;
      body_vector(f, a.psr_vectors + 3 * t, e, horizontal);
      if (decide) {
        const int j = sector_of(e[0], e[1]) * sectors.stride;
        if (e[2] < horizontal * static_cast<double>(sectors.dark[j])) continue;
        if (e[2] > horizontal * static_cast<double>(sectors.lit[j])) {
          best = INFINITY;
          break;
        }
      }
      double az, el;
      azimuth_elevation_of(e, horizontal, az, el);
      double m = el - horizon_at(hz, N, az);
      best = m > best ? m : best;
      if (mask_only && best > a.psr_threshold) break;
    }
    if (a.psr) a.psr[tile_pixels + pixel] = best > a.psr_threshold ? 0 : 255;
    if (a.psr_margin) a.psr_margin[tile_pixels + pixel] = static_cast<float>(best);
  }
}

}  // namespace lsh
