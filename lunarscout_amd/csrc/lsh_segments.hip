// lsh_segments.hip -- float64 ray-origin, direction, chord sampling and segment
// fitting of one horizon tile, one thread per (subpatch center, azimuth) job.
//
// Restates build_subpatch_segments_numba (geometry.py:575-673) and the Numba CPU
// job generate_segments_parallel/_generate_one (geometry_numba.py:172-364). All
// arithmetic is float64 ("projected-to-selenocentric ray-origin/curvature
// setup"); only the fitted 18-field segment is rounded to float32.
//
// Each job walks the ordered DEM list: DEM d+1 starts where DEM d's fitted
// segment ends, float32(s_end_km) * 1000, capped at DEM d's ray limit
// (geometry.py:659-663); a job whose start reached the maximum distance leaves
// the remaining DEM slots zero.

#include "lsh_internal.h"

namespace lsh {
namespace {

constexpr int kSamples = 16;
constexpr double kPi = 3.14159265358979323846;

struct Sample {
  double distance, column, row, latitude, longitude;
};

// _sample_chord (geometry_numba.py:173-222) without the unused terrain field.
__device__ bool sample_chord(const double* obs, const double* dir, double distance,
                             const SegmentDem& g, Sample& s) {
  double xv = obs[0] + dir[0] * distance;
  double yv = obs[1] + dir[1] * distance;
  double zv = obs[2] + dir[2] * distance;
  double lon = atan2(yv, xv);
  if (lon < 0.0) lon += 2.0 * kPi;
  double lat = atan2(zv, sqrt(xv * xv + yv * yv));
  double sl, cl;
  sincos(lat, &sl, &cl);
  double so, co;
  sincos(g.p[1], &so, &co);
  double delta = lon - g.p[2];
  double sd, cd;
  sincos(delta, &sd, &cd);
  double den = 1.0 + so * sl + co * cl * cd;
  if (fabs(den) < 1e-10) den = 1e-10;
  double scale = 2.0 * g.p[3] * g.p[0] / den;
  double px = scale * cl * sd + g.p[4];
  double py = scale * (co * sl - so * cl * cd) + g.p[5];
  double det = g.t[1] * g.t[5] - g.t[2] * g.t[4];
  double dx = px - g.t[0], dy = py - g.t[3];
  s.distance = distance;
  s.column = (g.t[5] * dx - g.t[2] * dy) / det;
  s.row = (-g.t[4] * dx + g.t[1] * dy) / det;
  s.latitude = lat;
  s.longitude = lon;
  return 0.0 <= s.column && s.column < g.width && 0.0 <= s.row && s.row < g.height;
}

// _build_samples (geometry_numba.py:226-297).
__device__ int build_samples(const double* obs, const double* dir, double start, double maximum,
                             const SegmentDem& g, Sample* samples) {
  Sample first;
  if (!sample_chord(obs, dir, start, g, first)) return 0;
  samples[0] = first;
  int count = 1;
  Sample final_inside = first, tmp;
  if (sample_chord(obs, dir, maximum, g, tmp)) {
    final_inside = tmp;
  } else {
    double low = start, high = maximum;
    Sample best = first;
    for (int i = 0; i < 24; ++i) {
      double middle = 0.5 * (low + high);
      if (sample_chord(obs, dir, middle, g, tmp)) {
        low = middle;
        best = tmp;
      } else {
        high = middle;
      }
    }
    final_inside = best;
  }
  double span = fmax(0.0, final_inside.distance - start);
  if (span > 1e-3) {
    for (int index = 1; index < 11; ++index) {
      double target = start + span * index / 10.0;
      if (target > final_inside.distance - 0.5) break;
      if (target - samples[count - 1].distance < 0.5) continue;
      if (sample_chord(obs, dir, target, g, tmp)) samples[count++] = tmp;
    }
    if (count < kSamples && final_inside.distance - samples[count - 1].distance > 0.5)
      samples[count++] = final_inside;
  }
  double required_end = samples[0].distance + 100.0;
  if (count < 4 || samples[count - 1].distance < required_end) {
    double step = fmax(10.0, g.map_resolution * 2.0);
    double current = samples[count - 1].distance;
    double extend_limit = fmax(final_inside.distance, required_end);
    while ((count < 4 || samples[count - 1].distance < required_end) && current < extend_limit &&
           count < kSamples) {
      current = fmin(extend_limit, current + step);
      bool inside = sample_chord(obs, dir, current, g, tmp);
      if (!inside || current - samples[count - 1].distance < 0.5) break;
      samples[count++] = tmp;
    }
  }
  return count;
}

// _solve (geometry_numba.py:30-63): Gauss-Jordan, partial pivoting, singular < 1e-12.
template <int n>
__device__ bool solve(double (&a)[n][n + 1], double* out) {
  for (int i = 0; i < n; ++i) {
    int pivot = i;
    double maximum = fabs(a[i][i]);
    for (int r = i + 1; r < n; ++r) {
      double v = fabs(a[r][i]);
      if (v > maximum) {
        maximum = v;
        pivot = r;
      }
    }
    if (maximum < 1e-12) return false;
    if (pivot != i)
      for (int c = 0; c <= n; ++c) {
        if (c < i) continue;
        double t = a[i][c];
        a[i][c] = a[pivot][c];
        a[pivot][c] = t;
      }
    double divisor = a[i][i];
    for (int c = 0; c <= n; ++c)
      if (c >= i) a[i][c] /= divisor;
    for (int r = 0; r < n; ++r) {
      if (r == i) continue;
      double f = a[r][i];
      if (fabs(f) < 1e-12) continue;
      for (int c = 0; c <= n; ++c)
        if (c >= i) a[r][c] -= f * a[i][c];
    }
  }
  for (int r = 0; r < n; ++r) out[r] = a[r][n];
  return true;
}

// _quartic (geometry_numba.py:67-89): no-intercept quartic in normalized distance,
// rescaled by span^-k.
__device__ void quartic(const Sample* s, int count, bool use_row, double v0, double anchor_km,
                        double span_km, double* out) {
  double a[4][5];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 5; ++c) a[r][c] = 0.0;
  for (int i = 0; i < count; ++i) {
    double p = (s[i].distance / 1000.0 - anchor_km) / span_km;
    double pw[4];
    pw[0] = p;
    for (int k = 1; k < 4; ++k) pw[k] = pw[k - 1] * p;
    double value = (use_row ? s[i].row : s[i].column) - v0;
    for (int r = 0; r < 4; ++r) {
      a[r][4] += pw[r] * value;
      for (int c = 0; c < 4; ++c) a[r][c] += pw[r] * pw[c];
    }
  }
  if (!solve<4>(a, out)) {
    for (int k = 0; k < 4; ++k) out[k] = 0.0;
    return;
  }
  double inv = 1.0 / span_km, scale = inv;
  for (int k = 0; k < 4; ++k) {
    out[k] *= scale;
    scale *= inv;
  }
}

// _chord_distance (geometry_numba.py:93-98).
__device__ double chord_distance(const double* obs, double lat, double lon, double radius) {
  double sl, cl, so, co;
  sincos(lat, &sl, &cl);
  sincos(lon, &so, &co);
  double x = radius * cl * co - obs[0];
  double y = radius * cl * so - obs[1];
  double z = radius * sl - obs[2];
  return sqrt(x * x + y * y + z * z);
}

// _fit_chord (geometry_numba.py:102-138): cubic planar -> chord correction.
__device__ void fit_chord(const Sample* s, int count, double res, const double* obs,
                          double radius, double* out) {
  out[0] = 1.0;
  out[1] = 0.0;
  out[2] = 0.0;
  if (count < 2 || res <= 0.0) return;
  double chord0 = chord_distance(obs, s[0].latitude, s[0].longitude, radius);
  double max_planar = 0.0;
  int used = 1;
  double a[3][4];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) a[r][c] = 0.0;
  // the point (0, 0) contributes nothing to the normal equations
  for (int i = 1; i < count; ++i) {
    if (s[i].column < 0.0 || s[i].row < 0.0) continue;
    double dx = (s[i].column - s[0].column) * res;
    double dy = (s[i].row - s[0].row) * res;
    double d = sqrt(dx * dx + dy * dy);
    double ch = chord_distance(obs, s[i].latitude, s[i].longitude, radius) - chord0;
    max_planar = fmax(max_planar, d);
    ++used;
    double pw[3] = {d, d * d, d * d * d};
    for (int r = 0; r < 3; ++r) {
      a[r][3] += pw[r] * ch;
      for (int c = 0; c < 3; ++c) a[r][c] += pw[r] * pw[c];
    }
  }
  if (used < 2 || max_planar < 500.0) return;
  double sol[3];
  if (!solve<3>(a, sol) || sol[0] < 0.5 || sol[0] > 2.0) return;
  out[0] = sol[0];
  out[1] = sol[1];
  out[2] = sol[2];
}

// _fit_one (geometry_numba.py:142-169) -> SEGMENT_FIELDS (contract.py:33-38).
__device__ void fit_one(const Sample* s, int count, double res, const double* obs,
                        double radius, double fallback_start, float* r) {
  for (int i = 0; i < kSegmentFields; ++i) r[i] = 0.0f;
  if (count < 3) {
    double x = count ? s[0].column : 0.0;
    double y = count ? s[0].row : 0.0;
    double start = count ? s[0].distance / 1000.0 : fallback_start / 1000.0;
    r[0] = r[2] = static_cast<float>(x);
    r[1] = r[3] = static_cast<float>(y);
    r[12] = r[13] = r[14] = static_cast<float>(start);
    r[15] = 1.0f;
    return;
  }
  double x0 = s[0].column, y0 = s[0].row;
  double anchor_km = s[0].distance / 1000.0;
  double end_km = s[count - 1].distance / 1000.0;
  double span_km = fmax(0.001, end_km - anchor_km);
  double xc[4], yc[4], ch[3];
  quartic(s, count, false, x0, anchor_km, span_km, xc);
  quartic(s, count, true, y0, anchor_km, span_km, yc);
  fit_chord(s, count, res, obs, radius, ch);
  r[0] = r[2] = static_cast<float>(x0);
  r[1] = r[3] = static_cast<float>(y0);
  for (int k = 0; k < 4; ++k) {
    r[4 + k] = static_cast<float>(xc[k]);
    r[8 + k] = static_cast<float>(yc[k]);
  }
  r[12] = r[14] = static_cast<float>(anchor_km);
  r[13] = static_cast<float>(end_km);
  for (int k = 0; k < 3; ++k) r[15 + k] = static_cast<float>(ch[k]);
}

__device__ int clamp_center(int requested, int dem_size, int sub) {
  int half = sub / 2;
  return min(max(requested, half), max(half, dem_size - half));
}

// One (center slot, azimuth) job over every DEM (geometry.py:605-663).
// `dems` holds the D DEMs of the tile (primary first); `out` the tile's
// [N][C][D][18] segment tensor.
__device__ __attribute__((always_inline)) void segment_job(
    const SegmentDem* dems, int D, float* out_tile, int N, int C, int count, int tile_column,
    int tile_row, int subpatch, double maximum_distance, double observer_elevation, int slot,
    int az) {
  const SegmentDem& primary = dems[0];
  int gr = slot / count, gc = slot % count;
  int cc = clamp_center(tile_column + (gc - 1) * subpatch + subpatch / 2,
                        primary.width, subpatch);
  int cr = clamp_center(tile_row + (gr - 1) * subpatch + subpatch / 2,
                        primary.height, subpatch);
  // inverse_stereographic (geometry.py:153-174) of the center pixel
  double x = primary.t[0] + primary.t[1] * cc + primary.t[2] * cr;
  double y = primary.t[3] + primary.t[4] * cc + primary.t[5] * cr;
  const double* p = primary.p;
  double xp = x - p[4], yp = y - p[5];
  double rho = hypot(xp, yp);
  double lat, lon;
  if (rho < 1e-9) {
    lat = p[1];
    lon = p[2];
  } else {
    double c = 2.0 * atan2(rho, 2.0 * p[3] * p[0]);
    double sc, ccos, so, co;
    sincos(c, &sc, &ccos);
    sincos(p[1], &so, &co);
    lat = asin(ccos * so + yp * sc * co / rho);
    lon = p[2] + atan2(xp * sc, rho * co * ccos - yp * so * sc);
  }
  // DemGrid.elevation at the clamped center (geometry.py:104-116, 608-610)
  double tc = fmin(fmax(static_cast<double>(cc), 0.0), primary.width - 1.001);
  double tr = fmin(fmax(static_cast<double>(cr), 0.0), primary.height - 1.001);
  int x1 = static_cast<int>(tc), y1 = static_cast<int>(tr);
  int x2 = min(x1 + 1, primary.width - 1), y2 = min(y1 + 1, primary.height - 1);
  const float* e = primary.elevation;
  double q11 = e[static_cast<int64_t>(y1) * primary.width + x1];
  double q12 = e[static_cast<int64_t>(y2) * primary.width + x1];
  double q21 = e[static_cast<int64_t>(y1) * primary.width + x2];
  double q22 = e[static_cast<int64_t>(y2) * primary.width + x2];
  double terrain = q11 * (x2 - tc) * (y2 - tr) + q21 * (tc - x1) * (y2 - tr) +
                   q12 * (x2 - tc) * (tr - y1) + q22 * (tc - x1) * (tr - y1);
  double correction_radius = p[0] + terrain;
  double radius = correction_radius + observer_elevation;
  double sl, cl, sn, cn;
  sincos(lat, &sl, &cl);
  sincos(lon, &sn, &cn);
  double obs[3] = {radius * cl * cn, radius * cl * sn, radius * sl};
  // azimuth_direction (geometry.py:228-234): clockwise from north
  double angle = kPi / 2.0 - az * 2.0 * kPi / N;
  double sa, ca;
  sincos(angle, &sa, &ca);
  double dir[3] = {ca * -sn + sa * (-sl * cn), ca * cn + sa * (-sl * sn), sa * cl};
  double norm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  for (int k = 0; k < 3; ++k) dir[k] /= norm;

  Sample samples[kSamples];
  double start = 1.0;
  float seg[kSegmentFields];
  for (int d = 0; d < D; ++d) {
    float* out = out_tile + ((static_cast<int64_t>(az) * C + slot) * D + d) * kSegmentFields;
    if (!(start < maximum_distance)) {
      for (int f = 0; f < kSegmentFields; ++f) out[f] = 0.0f;
      continue;
    }
    const SegmentDem& g = dems[d];
    int count = build_samples(obs, dir, start, g.ray_limit, g, samples);
    fit_one(samples, count, g.map_resolution, obs, correction_radius, start, seg);
    for (int f = 0; f < kSegmentFields; ++f) out[f] = seg[f];
    start = fmin(g.ray_limit, static_cast<double>(seg[13]) * 1000.0);
  }
}


__global__ void __launch_bounds__(128) segments_kernel(SegmentArgs args) {
  int64_t job = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (job >= static_cast<int64_t>(args.C) * args.N) return;
  segment_job(args.dem, args.D, args.out, args.N, args.C, args.count, args.tile_column,
              args.tile_row, args.subpatch, args.maximum_distance, args.observer_elevation,
              static_cast<int>(job / args.N), static_cast<int>(job % args.N));
}

// Every tile of a batch in one launch: blockIdx.y is the tile slot; the tile's
// site (its DEM list) and origin come from the device tables.
__global__ void __launch_bounds__(128) segments_batch_kernel(SegmentBatchArgs a) {
  const int tile = blockIdx.y;
  if (tile >= a.tile_count) return;
  int64_t job = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (job >= static_cast<int64_t>(a.C) * a.N) return;
  const lsh_batch_tile t = a.tiles[tile];
  const SegmentSite& site = a.sites[t.site];
  segment_job(site.dem, site.D, a.segments + tile * a.slot_floats, a.N, a.C, a.count,
              t.tile_column, t.tile_row, a.subpatch, a.maximum_distance,
              site.observer_elevation, static_cast<int>(job / a.N), static_cast<int>(job % a.N));
}

}  // namespace

hipError_t launch_segments(const SegmentArgs& args, hipStream_t stream) {
  int64_t jobs = static_cast<int64_t>(args.C) * args.N;
  int threads = 128;
  int blocks = static_cast<int>((jobs + threads - 1) / threads);
  hipLaunchKernelGGL(segments_kernel, dim3(blocks), dim3(threads), 0, stream, args);
  return hipGetLastError();
}

hipError_t launch_segments_batch(const SegmentBatchArgs& args, hipStream_t stream) {
  int64_t jobs = static_cast<int64_t>(args.C) * args.N;
  int threads = 128;
  int blocks = static_cast<int>((jobs + threads - 1) / threads);
  hipLaunchKernelGGL(segments_batch_kernel, dim3(blocks, args.tile_count), dim3(threads), 0,
                     stream, args);
  return hipGetLastError();
}

}  // namespace lsh
