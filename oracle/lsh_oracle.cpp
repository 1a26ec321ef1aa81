// lsh_oracle.cpp -- CPU oracle for the lunarscout horizon and lighting path.
//
// TEST INFRASTRUCTURE. This file restates the reference algorithm on the CPU so
// that the HIP product library can be checked against it. Only the tests, the
// smoke check and the CPU-baseline leg of the benchmark load it; the product
// library never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference repository root, under src/lunarscout/_numba_horizon/).
//
// Two arithmetic modes are provided for the march:
//   * method mode (double): the production schedule of hierarchy_maximum
//     evaluated entirely in float64 -- the parity target for the HIP kernel;
//   * faithful mode (float): the same schedule with the reference's float32
//     operation order, used to reproduce the captured C# traces. With the
//     diagnostic `cap` flag the level-0 advance is capped at the cell exit
//     (hierarchy.py:258-262), the schedule the captured traces encode.
//
// Build: see oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <utility>
#include <vector>

#include <omp.h>

extern "C" {

// One DEM raster with its affine transform and spherical stereographic
// projection parameters (geometry.py:24-116).
typedef struct {
  const float* elevation;  // float32[height][width], row-major
  int32_t width;
  int32_t height;
  double geo_transform[6];  // origin_x, t1, t2, origin_y, t4, t5
  double projection[6];     // radius, lat0, lon0, scale, false_e, false_n
} lsho_dem;

// A DEM together with its factor-four max pyramid (contract.py:274-344).
typedef struct {
  lsho_dem dem;
  const float* mips;       // concatenated levels 1..L-1
  const int32_t* levels;   // int32[L][4] (level, offset, width, height)
  int32_t level_count;
} lsho_pyramid;

}  // extern "C"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxSamples = 16;
constexpr int kSampleFields = 8;

inline bool valid_elevation(double v) {
  // cuda_backend.py:43-45 (finite and above -20000 m).
  return std::isfinite(v) && v > -20000.0;
}

// ---------------------------------------------------------------------------
// Pyramid (pyramid.py:19-31 layout, pyramid.py:110-138 / cuda_backend.py:134-153
// factor-four max over valid elevations, -32000 for an all-invalid block).
// ---------------------------------------------------------------------------

int pyramid_layout(int w, int h, int32_t* levels, int cap, int64_t* mip_count) {
  int count = 0;
  int64_t offset = 0;
  int cw = w, ch = h;
  while (true) {
    if (count >= cap) return -1;
    if (levels) {
      levels[count * 4 + 0] = count;
      levels[count * 4 + 1] = count == 0 ? 0 : static_cast<int32_t>(offset);
      levels[count * 4 + 2] = cw;
      levels[count * 4 + 3] = ch;
    }
    if (count > 0) offset += static_cast<int64_t>(cw) * ch;
    ++count;
    if (cw == 1 && ch == 1) break;
    cw = (cw + 3) / 4;
    ch = (ch + 3) / 4;
  }
  if (mip_count) *mip_count = offset;
  return count;
}

// ---------------------------------------------------------------------------
// FP64 host geometry (geometry.py:153-234, geometry_numba.py:172-297).
// ---------------------------------------------------------------------------

struct Geo {
  const float* elev;
  int w, h;
  double t[6];
  double p[6];
};

inline Geo make_geo(const lsho_dem& d) {
  Geo g;
  g.elev = d.elevation;
  g.w = d.width;
  g.h = d.height;
  for (int i = 0; i < 6; ++i) {
    g.t[i] = d.geo_transform[i];
    g.p[i] = d.projection[i];
  }
  return g;
}

// DemGrid.map_resolution_m (geometry.py:78-83).
inline double map_resolution(const double* t) {
  return (std::hypot(t[1], t[4]) + std::hypot(t[2], t[5])) * 0.5;
}

// DemGrid.elevation (geometry.py:104-116): bilinear with x2 = min(x1+1, w-1).
double grid_elevation(const Geo& g, double column, double row) {
  int x1 = static_cast<int>(column), y1 = static_cast<int>(row);
  int x2 = std::min(x1 + 1, g.w - 1), y2 = std::min(y1 + 1, g.h - 1);
  double q11 = g.elev[static_cast<int64_t>(y1) * g.w + x1];
  double q12 = g.elev[static_cast<int64_t>(y2) * g.w + x1];
  double q21 = g.elev[static_cast<int64_t>(y1) * g.w + x2];
  double q22 = g.elev[static_cast<int64_t>(y2) * g.w + x2];
  return q11 * (x2 - column) * (y2 - row) + q21 * (column - x1) * (y2 - row) +
         q12 * (x2 - column) * (row - y1) + q22 * (column - x1) * (row - y1);
}

// inverse_stereographic (geometry.py:153-174), rho threshold as given.
void inverse_stereographic(double x, double y, const double* p, double rho_eps,
                           double* lat, double* lon) {
  double xp = x - p[4], yp = y - p[5];
  double rho = std::hypot(xp, yp);
  if (rho < rho_eps) {
    *lat = p[1];
    *lon = p[2];
    return;
  }
  double c = 2.0 * std::atan2(rho, 2.0 * p[3] * p[0]);
  double sc = std::sin(c), cc = std::cos(c);
  double so = std::sin(p[1]), co = std::cos(p[1]);
  *lat = std::asin(cc * so + yp * sc * co / rho);
  *lon = p[2] + std::atan2(xp * sc, rho * co * cc - yp * so * sc);
}

// _sample_chord (geometry_numba.py:173-222).
bool sample_chord(const double* obs, const double* dir, double distance, const Geo& g,
                  double* s) {
  double xv = obs[0] + dir[0] * distance;
  double yv = obs[1] + dir[1] * distance;
  double zv = obs[2] + dir[2] * distance;
  double lon = std::atan2(yv, xv);
  if (lon < 0.0) lon += 2.0 * kPi;
  double lat = std::atan2(zv, std::sqrt(xv * xv + yv * yv));
  double sl = std::sin(lat), cl = std::cos(lat);
  double so = std::sin(g.p[1]), co = std::cos(g.p[1]);
  double delta = lon - g.p[2];
  double den = 1.0 + so * sl + co * cl * std::cos(delta);
  if (std::fabs(den) < 1e-10) den = 1e-10;
  double scale = 2.0 * g.p[3] * g.p[0] / den;
  double px = scale * cl * std::sin(delta) + g.p[4];
  double py = scale * (co * sl - so * cl * std::cos(delta)) + g.p[5];
  const double* t = g.t;
  double det = t[1] * t[5] - t[2] * t[4];
  double dx = px - t[0], dy = py - t[3];
  double column = (t[5] * dx - t[2] * dy) / det;
  double row = (-t[4] * dx + t[1] * dy) / det;
  bool inside = 0.0 <= column && column < g.w && 0.0 <= row && row < g.h;
  double terrain = 0.0;
  if (inside) {
    int x1 = static_cast<int>(column), y1 = static_cast<int>(row);
    int x2 = std::min(x1 + 1, g.w - 1), y2 = std::min(y1 + 1, g.h - 1);
    auto E = [&](int yy, int xx) {
      return static_cast<double>(g.elev[static_cast<int64_t>(yy) * g.w + xx]);
    };
    terrain = E(y1, x1) * (x2 - column) * (y2 - row) + E(y1, x2) * (column - x1) * (y2 - row) +
              E(y2, x1) * (x2 - column) * (row - y1) + E(y2, x2) * (column - x1) * (row - y1);
  }
  s[0] = distance;
  s[1] = column;
  s[2] = row;
  s[3] = lat;
  s[4] = lon;
  s[5] = row;
  s[6] = column;
  s[7] = terrain;
  return inside;
}

// _build_samples (geometry_numba.py:226-297).
int build_samples(const double* obs, const double* dir, double start, double maximum,
                  const Geo& g, double* samples /*[16][8]*/) {
  std::memset(samples, 0, sizeof(double) * kMaxSamples * kSampleFields);
  double first[kSampleFields];
  if (!sample_chord(obs, dir, start, g, first)) return 0;
  std::memcpy(samples, first, sizeof(first));
  int count = 1;
  double final_inside[kSampleFields];
  std::memcpy(final_inside, first, sizeof(first));
  double tmp[kSampleFields];
  if (sample_chord(obs, dir, maximum, g, tmp)) {
    std::memcpy(final_inside, tmp, sizeof(tmp));
  } else {
    double low = start, high = maximum;
    double best[kSampleFields];
    std::memcpy(best, first, sizeof(first));
    for (int i = 0; i < 24; ++i) {
      double middle = 0.5 * (low + high);
      if (sample_chord(obs, dir, middle, g, tmp)) {
        low = middle;
        std::memcpy(best, tmp, sizeof(tmp));
      } else {
        high = middle;
      }
    }
    std::memcpy(final_inside, best, sizeof(best));
  }
  double span = std::max(0.0, final_inside[0] - start);
  if (span > 1e-3) {
    for (int index = 1; index < 11; ++index) {
      double target = start + span * index / 10.0;
      if (target > final_inside[0] - 0.5) break;
      if (target - samples[(count - 1) * kSampleFields] < 0.5) continue;
      if (sample_chord(obs, dir, target, g, tmp)) {
        std::memcpy(samples + count * kSampleFields, tmp, sizeof(tmp));
        ++count;
      }
    }
    if (count < kMaxSamples && final_inside[0] - samples[(count - 1) * kSampleFields] > 0.5) {
      std::memcpy(samples + count * kSampleFields, final_inside, sizeof(final_inside));
      ++count;
    }
  }
  double required_end = samples[0] + 100.0;
  double res = 0.5 * (std::sqrt(g.t[1] * g.t[1] + g.t[4] * g.t[4]) +
                      std::sqrt(g.t[2] * g.t[2] + g.t[5] * g.t[5]));
  if (count < 4 || samples[(count - 1) * kSampleFields] < required_end) {
    double step = std::max(10.0, res * 2.0);
    double current = samples[(count - 1) * kSampleFields];
    double extend_limit = std::max(final_inside[0], required_end);
    while ((count < 4 || samples[(count - 1) * kSampleFields] < required_end) &&
           current < extend_limit && count < kMaxSamples) {
      current = std::min(extend_limit, current + step);
      bool inside = sample_chord(obs, dir, current, g, tmp);
      if (!inside || current - samples[(count - 1) * kSampleFields] < 0.5) break;
      std::memcpy(samples + count * kSampleFields, tmp, sizeof(tmp));
      ++count;
    }
  }
  return count;
}

// _solve (geometry_numba.py:30-63): Gauss-Jordan with partial pivoting.
bool solve(int n, double* m /*[n][n]*/, double* rhs, double* out) {
  double a[4][5];
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < n; ++c) a[r][c] = m[r * n + c];
    a[r][n] = rhs[r];
  }
  for (int i = 0; i < n; ++i) {
    int pivot = i;
    double maximum = std::fabs(a[i][i]);
    for (int r = i + 1; r < n; ++r) {
      double v = std::fabs(a[r][i]);
      if (v > maximum) {
        maximum = v;
        pivot = r;
      }
    }
    if (maximum < 1e-12) {
      for (int r = 0; r < n; ++r) out[r] = 0.0;
      return false;
    }
    if (pivot != i)
      for (int c = i; c <= n; ++c) std::swap(a[i][c], a[pivot][c]);
    double divisor = a[i][i];
    for (int c = i; c <= n; ++c) a[i][c] /= divisor;
    for (int r = 0; r < n; ++r) {
      if (r == i) continue;
      double f = a[r][i];
      if (std::fabs(f) < 1e-12) continue;
      for (int c = i; c <= n; ++c) a[r][c] -= f * a[i][c];
    }
  }
  for (int r = 0; r < n; ++r) out[r] = a[r][n];
  return true;
}

// _quartic (geometry_numba.py:67-89).
void quartic(const double* samples, int count, int column, double v0, double anchor_km,
             double span_km, double* out) {
  double normal[16] = {0}, rhs[4] = {0};
  for (int i = 0; i < count; ++i) {
    double p = (samples[i * kSampleFields] / 1000.0 - anchor_km) / span_km;
    double pw[4];
    pw[0] = p;
    for (int k = 1; k < 4; ++k) pw[k] = pw[k - 1] * p;
    double value = samples[i * kSampleFields + column] - v0;
    for (int r = 0; r < 4; ++r) {
      rhs[r] += pw[r] * value;
      for (int c = 0; c < 4; ++c) normal[r * 4 + c] += pw[r] * pw[c];
    }
  }
  if (!solve(4, normal, rhs, out)) {
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
double chord_distance(const double* obs, double lat, double lon, double radius) {
  double cl = std::cos(lat);
  double x = radius * cl * std::cos(lon) - obs[0];
  double y = radius * cl * std::sin(lon) - obs[1];
  double z = radius * std::sin(lat) - obs[2];
  return std::sqrt(x * x + y * y + z * z);
}

// _fit_chord (geometry_numba.py:102-138).
void fit_chord(const double* samples, int count, double res, const double* obs,
               double radius, double* out) {
  out[0] = 1.0;
  out[1] = 0.0;
  out[2] = 0.0;
  if (count < 2 || res <= 0.0) return;
  double planar[16] = {0}, chord[16] = {0};
  double chord0 = chord_distance(obs, samples[3], samples[4], radius);
  double max_planar = 0.0;
  int used = 1;
  for (int i = 1; i < std::min(count, 16); ++i) {
    const double* s = samples + i * kSampleFields;
    if (s[1] < 0.0 || s[2] < 0.0) continue;
    double dx = (s[1] - samples[1]) * res;
    double dy = (s[2] - samples[2]) * res;
    double d = std::sqrt(dx * dx + dy * dy);
    planar[used] = d;
    chord[used] = chord_distance(obs, s[3], s[4], radius) - chord0;
    max_planar = std::max(max_planar, d);
    ++used;
  }
  if (used < 2 || max_planar < 500.0) return;
  double normal[9] = {0}, rhs[3] = {0};
  for (int i = 0; i < used; ++i) {
    double x = planar[i];
    double pw[3] = {x, x * x, x * x * x};
    for (int r = 0; r < 3; ++r) {
      rhs[r] += pw[r] * chord[i];
      for (int c = 0; c < 3; ++c) normal[r * 3 + c] += pw[r] * pw[c];
    }
  }
  double sol[3];
  bool ok = solve(3, normal, rhs, sol);
  if (!ok || sol[0] < 0.5 || sol[0] > 2.0) return;
  out[0] = sol[0];
  out[1] = sol[1];
  out[2] = sol[2];
}

// _fit_one (geometry_numba.py:142-169) -> float32[18] segment (contract.py:33-38).
void fit_one(const double* samples, int count, double res, const double* obs, double radius,
             double fallback_start, float* r) {
  for (int i = 0; i < 18; ++i) r[i] = 0.0f;
  if (count < 3) {
    double x = count ? samples[1] : 0.0;
    double y = count ? samples[2] : 0.0;
    double start = count ? samples[0] / 1000.0 : fallback_start / 1000.0;
    r[0] = static_cast<float>(x);
    r[1] = static_cast<float>(y);
    r[2] = static_cast<float>(x);
    r[3] = static_cast<float>(y);
    r[12] = static_cast<float>(start);
    r[13] = static_cast<float>(start);
    r[14] = static_cast<float>(start);
    r[15] = 1.0f;
    return;
  }
  double x0 = samples[1], y0 = samples[2];
  double anchor_km = samples[0] / 1000.0;
  double end_km = samples[(count - 1) * kSampleFields] / 1000.0;
  double span_km = std::max(0.001, end_km - anchor_km);
  double xc[4], yc[4], ch[3];
  quartic(samples, count, 1, x0, anchor_km, span_km, xc);
  quartic(samples, count, 2, y0, anchor_km, span_km, yc);
  fit_chord(samples, count, res, obs, radius, ch);
  r[0] = static_cast<float>(x0);
  r[1] = static_cast<float>(y0);
  r[2] = static_cast<float>(x0);
  r[3] = static_cast<float>(y0);
  for (int k = 0; k < 4; ++k) {
    r[4 + k] = static_cast<float>(xc[k]);
    r[8 + k] = static_cast<float>(yc[k]);
  }
  r[12] = static_cast<float>(anchor_km);
  r[13] = static_cast<float>(end_km);
  r[14] = static_cast<float>(anchor_km);
  for (int k = 0; k < 3; ++k) r[15 + k] = static_cast<float>(ch[k]);
}

// clamp_subpatch_center (geometry.py:465-467).
inline int clamp_center(int requested, int dem_size, int sub) {
  int half = sub / 2;
  return std::min(std::max(requested, half), std::max(half, dem_size - half));
}

// Observer frame of one subpatch center (geometry.py:605-620).
void center_frame(const Geo& primary, int cc, int cr, int N, double obs_h, double* observer,
                  double* correction_radius, double* directions /*[N][3]*/) {
  double x = primary.t[0] + primary.t[1] * cc + primary.t[2] * cr;
  double y = primary.t[3] + primary.t[4] * cc + primary.t[5] * cr;
  double lat, lon;
  inverse_stereographic(x, y, primary.p, 1e-9, &lat, &lon);
  double tc = std::min(std::max(static_cast<double>(cc), 0.0), primary.w - 1.001);
  double tr = std::min(std::max(static_cast<double>(cr), 0.0), primary.h - 1.001);
  double terrain = grid_elevation(primary, tc, tr);
  *correction_radius = primary.p[0] + terrain;
  double radius = *correction_radius + obs_h;
  double cl = std::cos(lat), sl = std::sin(lat), co = std::cos(lon), so = std::sin(lon);
  observer[0] = radius * cl * co;
  observer[1] = radius * cl * so;
  observer[2] = radius * sl;
  double east[3] = {-so, co, 0.0};
  double north[3] = {-sl * co, -sl * so, cl};
  for (int a = 0; a < N; ++a) {
    double az = a * 2.0 * kPi / N;
    double angle = kPi / 2.0 - az;
    double ca = std::cos(angle), sa = std::sin(angle);
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = ca * east[k] + sa * north[k];
    double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) directions[a * 3 + k] = d[k] / n;
  }
}

// ---------------------------------------------------------------------------
// The production march: hierarchy_maximum (cuda_backend.py:656-874) with the
// tile-level subpatch interpolation (cuda_backend.py:1007-1134).
// ---------------------------------------------------------------------------

struct MarchDem {
  const float* level0;
  const float* mips;
  const int32_t* levels;
  int L;
  int w, h;
};

struct MarchOpts {
  bool cap;            // diagnostic cell-exit cap (hierarchy.py:258-262)
  bool stable_slope;   // (h - oz)(2R + h + oz) form of the level-0 slope
  float* trace;        // nullable, [trace_cap][12]
  int trace_cap;
  int trace_count;
  int64_t samples;     // level-0 samples (action 4)
  int64_t tests;       // hierarchy bound tests (actions 0, 1, 2)
  int64_t steps;       // every trace row
  // Frozen-schedule evaluation (the adjoint's convention): when eval_level0 is
  // set, every level-0 sample the march takes on its own DEM is also evaluated on
  // eval_level0 (same grid) against eval_observer_z, and eval_current keeps the
  // maximum of those slopes; the schedule itself is the march's own.
  const float* eval_level0 = nullptr;
  double eval_observer_z = 0.0;
  double eval_current = -1e30;
};

inline float cell_value(const MarchDem& d, int level, int x, int y) {
  if (level == 0) return d.level0[static_cast<int64_t>(y) * d.w + x];
  const int32_t* lv = d.levels + level * 4;
  return d.mips[lv[1] + static_cast<int64_t>(y) * lv[2] + x];
}

// hierarchy_bilinear_bound (cuda_backend.py:47-60).
template <class T>
T bilinear_bound(const MarchDem& d, int level, int cx, int cy) {
  const int32_t* lv = d.levels + level * 4;
  int lw = lv[2], lh = lv[3];
  float maximum = -32000.0f;
  for (int ny = cy; ny < std::min(cy + 2, lh); ++ny)
    for (int nx = cx; nx < std::min(cx + 2, lw); ++nx) {
      float v = cell_value(d, level, nx, ny);
      if (valid_elevation(v) && v > maximum) maximum = v;
    }
  return static_cast<T>(maximum);
}

// bilinear (cuda_backend.py:107-132): clamp to size - 1.0001, any invalid tap -> -32000.
template <class T>
T bilinear(const float* e, int w, int h, T column, T row) {
  column = std::min(std::max(column, T(0)), T(w) - T(1.0001));
  row = std::min(std::max(row, T(0)), T(h) - T(1.0001));
  int x0 = static_cast<int>(std::floor(column));
  int y0 = static_cast<int>(std::floor(row));
  int x1 = std::min(x0 + 1, w - 1), y1 = std::min(y0 + 1, h - 1);
  float q00 = e[static_cast<int64_t>(y0) * w + x0];
  float q10 = e[static_cast<int64_t>(y0) * w + x1];
  float q01 = e[static_cast<int64_t>(y1) * w + x0];
  float q11 = e[static_cast<int64_t>(y1) * w + x1];
  if (!(valid_elevation(q00) && valid_elevation(q10) && valid_elevation(q01) &&
        valid_elevation(q11)))
    return T(-32000.0);
  T tx = column - T(x0);
  T ty = row - T(y0);
  T top = T(q00) + tx * (T(q10) - T(q00));
  T bottom = T(q01) + tx * (T(q11) - T(q01));
  return top + ty * (bottom - top);
}

// evaluate / tangent / planar_chord (cuda_backend.py:62-93), same operation order.
template <class T>
inline T eval_quartic(T x0, T a1, T a2, T a3, T a4, T d) {
  T d2 = d * d, d3 = d2 * d, d4 = d2 * d2;
  T r = x0 + a1 * d;
  r = r + a2 * d2;
  r = r + a3 * d3;
  return r + a4 * d4;
}
template <class T>
inline T eval_tangent(T a1, T a2, T a3, T a4, T d) {
  T d2 = d * d, d3 = d2 * d;
  T r = a1 + T(2) * a2 * d;
  r = r + T(3) * a3 * d2;
  return r + T(4) * a4 * d3;
}

template <class T>
void trace_row(MarchOpts& o, T s, T true_m, int level, int cx, int cy, T px, T py, T maxh,
               T sh, T slope, T advance, int action) {
  ++o.steps;
  if (action == 4) ++o.samples;
  if (action == 0 || action == 1 || action == 2) ++o.tests;
  if (!o.trace || o.trace_count >= o.trace_cap) {
    if (o.trace) ++o.trace_count;
    return;
  }
  float* r = o.trace + o.trace_count * 12;
  r[0] = float(s);
  r[1] = float(true_m);
  r[2] = float(level);
  r[3] = float(cx);
  r[4] = float(cy);
  r[5] = float(px);
  r[6] = float(py);
  r[7] = float(maxh);
  r[8] = float(sh);
  r[9] = float(slope);
  r[10] = float(advance);
  r[11] = float(action);
  ++o.trace_count;
}

// hierarchy_maximum (cuda_backend.py:656-874). Segment fields follow SEGMENT_FIELDS
// (contract.py:33-38): [0,1] start px, [4..7] a1..a4, [8..11] b1..b4, [12] s_start,
// [13] s_end, [14] s_start_chord, [15..17] chord c1..c3.
template <class T>
T hierarchy_maximum(const T* sg, const MarchDem& d, T observer_z, T radius, T res,
                    int pass_index, T initial, MarchOpts& o) {
  const T nan = std::numeric_limits<T>::quiet_NaN();
  T start_x = sg[0], start_y = sg[1];
  T a1 = sg[4], a2 = sg[5], a3 = sg[6], a4 = sg[7];
  T b1 = sg[8], b2 = sg[9], b3 = sg[10], b4 = sg[11];
  T s_start = sg[12], s_end = sg[13], s_start_chord = sg[14];
  T c1 = sg[15], c2 = sg[16], c3 = sg[17];
  T s = std::max(s_start, T(0.001));
  T current = std::isinf(initial) ? T(-1e30) : initial;
  T minimum_step = (T(0.5) * res) * T(0.001);
  T primary_far_step = (T(0.8) * res) * T(0.001);
  bool ray_out = false;
  while (s <= s_end) {
    T delta = s - s_start;
    T px = eval_quartic(start_x, a1, a2, a3, a4, delta);
    T py = eval_quartic(start_y, b1, b2, b3, b4, delta);
    if (std::isnan(px) || std::isnan(py) || px < T(0) || py < T(0) || px >= T(d.w - 1) ||
        py >= T(d.h - 1))
      break;
    T plx = (px - start_x) * res;
    T ply = (py - start_y) * res;
    T planar_m = std::sqrt(plx * plx + ply * ply);
    T true_m;
    if (s < T(0.5)) {
      true_m = s * T(1000);
    } else {
      T p2 = planar_m * planar_m, p3 = p2 * planar_m;
      T chord = c1 * planar_m;
      chord = chord + c2 * p2;
      chord = chord + c3 * p3;
      true_m = s_start_chord * T(1000) + chord;
    }
    T footprint = true_m * T(0.004363323129985824);
    int level = d.L - 1;
    while (level > 0) {
      T side = T(static_cast<double>(int64_t(1) << (level * 2))) * res;
      if (side <= footprint) break;
      --level;
    }
    while (level >= 0) {
      const int32_t* lv = d.levels + level * 4;
      int lw = lv[2], lh = lv[3];
      int shift = level * 2;
      int scale = 1 << shift;
      int cx = static_cast<int>(px) >> shift;
      int cy = static_cast<int>(py) >> shift;
      if (!(0 <= cx && cx < lw && 0 <= cy && cy < lh)) {
        T step_km = (T(scale) * res) * T(0.001);
        T advance = std::max(T(0.001), step_km);
        trace_row<T>(o, s, true_m, level, cx, cy, px, py, nan, nan, nan, advance, 3);
        s = s + advance;
        ray_out = true;
        break;
      }
      T maxh = bilinear_bound<T>(d, level, cx, cy);
      T min_x = T(cx * scale), min_y = T(cy * scale);
      T max_x = min_x + T(scale), max_y = min_y + T(scale);
      T dx = eval_tangent(a1, a2, a3, a4, delta);
      T dy = eval_tangent(b1, b2, b3, b4, delta);
      T inv_x = std::fabs(dx) > T(1e-8) ? T(1) / dx : T(1e30);
      T inv_y = std::fabs(dy) > T(1e-8) ? T(1) / dy : T(1e30);
      T t1 = (min_x - px) * inv_x, t2 = (max_x - px) * inv_x;
      T t3 = (min_y - py) * inv_y, t4 = (max_y - py) * inv_y;
      T exit_distance = std::min(std::max(t1, t2), std::max(t3, t4));
      T fallback = ((T(scale) * res) * T(0.5)) * T(0.001);
      T to_exit = exit_distance > T(0) ? exit_distance : fallback;
      if (maxh < T(-20000)) {
        T advance = to_exit > T(0) ? to_exit + T(0.0001) : fallback;
        trace_row<T>(o, s, true_m, level, cx, cy, px, py, maxh, nan, nan, advance, 2);
        s = s + advance;
        break;
      }
      T block_m = T(scale) * res;
      T true_near = std::max(true_m - block_m, T(1));
      T observer_radius = radius + observer_z;
      T possible;
      if (s < T(0.5)) {
        possible = (maxh - observer_z) / true_near;
      } else {
        T sq = true_near * true_near;
        T hd = maxh - observer_z;
        T rs = T(2) * radius + maxh;
        rs = rs + observer_z;
        T num = hd * rs - sq;
        T den = T(2) * observer_radius;
        T lz = num / den;
        T lxx = sq - lz * lz;
        T lx = lxx > T(0) ? std::sqrt(lxx) : T(1e-6);
        possible = lz / lx;
      }
      if (possible <= current) {
        T advance = to_exit > T(0) ? to_exit + T(0.0001) : fallback;
        trace_row<T>(o, s, true_m, level, cx, cy, px, py, maxh, nan, nan, advance, 1);
        s = s + advance;
        break;
      }
      if (level == 0) {
        T height = bilinear<T>(d.level0, d.w, d.h, px, py);
        T slope = T(-1e30);
        if (valid_elevation(double(height))) {
          if (s < T(0.5)) {
            slope = (height - observer_z) / true_m;
          } else {
            T sq = true_m * true_m;
            T num;
            if (o.stable_slope) {
              T hd = height - observer_z;
              T rs = T(2) * radius + height;
              rs = rs + observer_z;
              num = hd * rs - sq;
            } else {
              T pr = radius + height;
              T rd = pr - observer_radius;
              T rsum = pr + observer_radius;
              num = rd * rsum - sq;
            }
            T den = T(2) * observer_radius;
            T lz = num / den;
            T lxx = sq - lz * lz;
            T lx = lxx > T(0) ? std::sqrt(lxx) : T(1e-6);
            slope = lz / lx;
          }
          if (slope > current) current = slope;
        }
        if (o.eval_level0) {
          T he = bilinear<T>(o.eval_level0, d.w, d.h, px, py);
          if (valid_elevation(double(he))) {
            T ez = T(o.eval_observer_z);
            T es;
            if (s < T(0.5)) {
              es = (he - ez) / true_m;
            } else {
              T sq = true_m * true_m;
              T hd = he - ez;
              T rs = T(2) * radius + he;
              rs = rs + ez;
              T lz = (hd * rs - sq) / (T(2) * (radius + ez));
              T lxx = sq - lz * lz;
              es = lz / (lxx > T(0) ? std::sqrt(lxx) : T(1e-6));
            }
            if (double(es) > o.eval_current) o.eval_current = double(es);
          }
        }
        T mag = std::sqrt(dx * dx + dy * dy);
        T pixel_step = mag > T(1e-6) ? T(1) / mag : T(0.0005);
        T margin = current - slope;
        T margin_step = margin > T(0) ? ((margin * true_m) * T(1.732)) / T(1000) : T(0);
        T angular_step = (true_m * T(0.00151)) / T(1000);
        T advance = std::max(pixel_step, std::min(margin_step, angular_step));
        if (s < T(0.5)) advance = advance * T(0.25);
        T floor_step =
            (pass_index == 0 && true_m >= T(100)) ? primary_far_step : minimum_step;
        advance = std::max(advance, floor_step);
        if (o.cap) {
          T boundary = to_exit > T(0) ? to_exit + T(0.000001) : fallback;
          advance = std::min(advance, boundary);
        }
        trace_row<T>(o, s, true_m, level, cx, cy, px, py, maxh, height, slope, advance, 4);
        s = s + advance;
        break;
      }
      trace_row<T>(o, s, true_m, level, cx, cy, px, py, maxh, nan, nan, T(0), 0);
      --level;
    }
    if (ray_out) break;
  }
  return current > T(-1e29) ? current : -std::numeric_limits<T>::infinity();
}

// Resolution, radius and scale inputs the kernel derives from its map/projection
// parameters (cuda_backend.py:1032-1039, 1069-1073, 1124-1127).
struct PassGeometry {
  double column_res, row_res;  // hypot(t1, t4), hypot(t2, t5)
  double radius;
};

inline PassGeometry pass_geometry(const lsho_dem& d, bool f32) {
  PassGeometry g;
  if (f32) {
    float t1 = float(d.geo_transform[1]), t4 = float(d.geo_transform[4]);
    float t2 = float(d.geo_transform[2]), t5 = float(d.geo_transform[5]);
    g.column_res = std::sqrt(t1 * t1 + t4 * t4);
    g.row_res = std::sqrt(t2 * t2 + t5 * t5);
    g.radius = float(d.projection[0]);
  } else {
    const double* t = d.geo_transform;
    g.column_res = std::sqrt(t[1] * t[1] + t[4] * t[4]);
    g.row_res = std::sqrt(t[2] * t[2] + t[5] * t[5]);
    g.radius = d.projection[0];
  }
  return g;
}

// One (pixel, azimuth) ray over every DEM pass (cuda_backend.py:1014-1134 with the
// running maximum carried across passes, cuda_backend.py:1445-1552).
template <class T>
T ray_all_passes(const lsho_pyramid* pyrs, int D, const float* segments, int N, int C,
                 int az, int column, int row, int tile_col, int tile_row, int tile_w,
                 int subpatch, double obs_h, MarchOpts& o,
                 const float* const* eval_level0 = nullptr) {
  const lsho_dem& primary = pyrs[0].dem;
  bool f32 = sizeof(T) == sizeof(float);
  int count = tile_w / subpatch + 2;
  float gxf = float(column - subpatch / 2) / float(subpatch) + 1.0f;
  float gyf = float(row - subpatch / 2) / float(subpatch) + 1.0f;
  int left = int(gxf), top = int(gyf);
  float tx = gxf - float(left), ty = gyf - float(top);
  if (left < 0) { left = 0; tx = 0.0f; }
  if (top < 0) { top = 0; ty = 0.0f; }
  if (left > count - 2) { left = count - 2; tx = 1.0f; }
  if (top > count - 2) { top = count - 2; ty = 1.0f; }
  int i00 = top * count + left, i10 = i00 + 1;
  int i01 = (top + 1) * count + left, i11 = i01 + 1;
  int rl = tile_col + (left - 1) * subpatch + subpatch / 2, rr = rl + subpatch;
  int rt = tile_row + (top - 1) * subpatch + subpatch / 2, rb = rt + subpatch;
  int cl = clamp_center(rl, primary.width, subpatch) - tile_col;
  int cr = clamp_center(rr, primary.width, subpatch) - tile_col;
  int ct = clamp_center(rt, primary.height, subpatch) - tile_row;
  int cb = clamp_center(rb, primary.height, subpatch) - tile_row;
  PassGeometry pg0 = pass_geometry(primary, f32);
  T obs_terrain = bilinear<T>(primary.elevation, primary.width, primary.height,
                              T(tile_col + column), T(tile_row + row));
  T observer_z = obs_terrain + T(f32 ? double(float(obs_h)) : obs_h);
  if (eval_level0) {
    o.eval_observer_z = double(bilinear<T>(eval_level0[0], primary.width, primary.height,
                                           T(tile_col + column), T(tile_row + row))) + obs_h;
    o.eval_current = -1e30;
  }
  T running = -std::numeric_limits<T>::infinity();
  for (int p = 0; p < D; ++p) {
    o.eval_level0 = eval_level0 ? eval_level0[p] : nullptr;
    PassGeometry pg = pass_geometry(pyrs[p].dem, f32);
    T ratio = T(pg0.column_res) / T(pg.column_res);
    T sxl = T(column - cl) * ratio, sxr = T(column - cr) * ratio;
    T syt = T(row - ct) * ratio, syb = T(row - cb) * ratio;
    T sg[18];
    const float* base = segments + (static_cast<int64_t>(az) * C) * D * 18;
    auto at = [&](int idx, int f) { return T(base[(static_cast<int64_t>(idx) * D + p) * 18 + f]); };
    for (int f = 0; f < 18; ++f) {
      if (f == 2 || f == 3) { sg[f] = 0; continue; }
      T v00 = at(i00, f), v10 = at(i10, f), v01 = at(i01, f), v11 = at(i11, f);
      if (f == 0) { v00 = v00 + sxl; v10 = v10 + sxr; v01 = v01 + sxl; v11 = v11 + sxr; }
      if (f == 1) { v00 = v00 + syt; v10 = v10 + syt; v01 = v01 + syb; v11 = v11 + syb; }
      T upper, lower, value;
      if (f32) {
        upper = T(std::fma(float(v10 - v00), tx, float(v00)));
        lower = T(std::fma(float(v11 - v01), tx, float(v01)));
        value = T(std::fma(float(lower - upper), ty, float(upper)));
      } else {
        upper = (v10 - v00) * T(tx) + v00;
        lower = (v11 - v01) * T(tx) + v01;
        value = (lower - upper) * T(ty) + upper;
      }
      sg[f] = value;
    }
    T res = (T(pg.column_res) + T(pg.row_res)) * T(0.5);
    MarchDem md{pyrs[p].dem.elevation, pyrs[p].mips, pyrs[p].levels, pyrs[p].level_count,
                pyrs[p].dem.width, pyrs[p].dem.height};
    running = hierarchy_maximum<T>(sg, md, observer_z, T(pg.radius), res, p, running, o);
  }
  return running;
}

// ---------------------------------------------------------------------------
// Lighting (lightmap.py:15-73, 116-121; psr.py:17-140; lightmap_cpu.py:25-183).
// ---------------------------------------------------------------------------

// _HALF_CIRCLE and _MAX_PHOTONS (lightmap.py:15-29), built in float64 then rounded.
struct Disk {
  float half[16];
  float max_photons;
  Disk() {
    float sum = 0.0f;
    for (int i = 0; i < 16; ++i) {
      half[i] = float(std::sqrt(64.0 - (7.5 - i) * (7.5 - i)) / 8.0 * double(0.27f));
      sum = sum + half[i];
    }
    max_photons = 2.0f * sum;
  }
};
const Disk kDisk;

// _sun_fraction_reference (lightmap.py:32-73): C# BuilderSunFraction. T = float
// reproduces the float32 accumulation; T = double evaluates the same slices in float64.
//
// The reference walks 1440 bins: x4.0 bins per degree, half a bin = 0.125 deg,
// slices 0.135 bins apart. Another N scales the three to its bin width (equal
// to the reference's constants, bit for bit, at N = 1440).
template <class T>
T sun_fraction(const float* hz, int N, T az_deg, T el_deg) {
  const float bins_per_degree = float(N) / 360.0f;
  const T per_degree = T(bins_per_degree);
  const T half_bucket = T(0.5f / bins_per_degree);
  const T step = T(0.135f * (bins_per_degree * 0.25f));
  T left_pos = (az_deg - T(0.27f) - half_bucket) * per_degree;
  int left = static_cast<int>(left_pos);  // truncation toward zero
  T frac = left_pos - T(left);
  if (left < 0) left += N;
  else if (left >= N) left -= N;
  int right = left + 1;
  if (right >= N) right = 0;
  T le = hz[left], re = hz[right];
  T delta = re - le;
  T photons = 0;
  for (int i = 0; i < 16; ++i) {
    T col = T(kDisk.half[i]);
    T horizon = frac * delta + le;
    T top = el_deg + col;
    if (horizon < top) {
      T ad = top - horizon;
      T ch = col + col;
      if (ad > ch) ad = ch;
      photons = photons + ad;
    }
    frac = frac + step;
    if (frac >= T(1)) {
      left = right;
      right += 1;
      if (right >= N) right = 0;
      le = re;
      re = hz[right];
      delta = re - le;
      frac = frac - T(1);
    }
  }
  return photons / T(kDisk.max_photons);
}

// _pixel_frame (psr.py:29-47): rotation columns are east, north, up.
void pixel_frame(const lsho_dem& d, int row, int column, double rot[9], double tr[3]) {
  const double* t = d.geo_transform;
  double x = t[0] + t[1] * column + t[2] * row;
  double y = t[3] + t[4] * column + t[5] * row;
  double lat, lon;
  inverse_stereographic(x, y, d.projection, 1e-9, &lat, &lon);
  double cl = std::cos(lat), sl = std::sin(lat), co = std::cos(lon), so = std::sin(lon);
  double up[3] = {cl * co, cl * so, sl};
  double east[3] = {-so, co, 0.0};
  double north[3] = {-sl * co, -sl * so, cl};
  double radius = d.projection[0] + double(d.elevation[static_cast<int64_t>(row) * d.width + column]);
  double obs[3] = {radius * up[0], radius * up[1], radius * up[2]};
  // rotation = stack((east, north, up)).T : rot[k][j] = basis_j[k]
  for (int k = 0; k < 3; ++k) {
    rot[k * 3 + 0] = east[k];
    rot[k * 3 + 1] = north[k];
    rot[k * 3 + 2] = up[k];
  }
  for (int j = 0; j < 3; ++j) {
    double acc = 0.0;
    for (int k = 0; k < 3; ++k) acc += obs[k] * rot[k * 3 + j];
    tr[j] = -acc;
  }
}

// _azimuth_elevation_deg (psr.py:50-58).
inline void azimuth_elevation(const double* v, const double* rot, const double* tr,
                              double* az, double* el) {
  double e[3];
  for (int j = 0; j < 3; ++j)
    e[j] = v[0] * rot[0 * 3 + j] + v[1] * rot[1 * 3 + j] + v[2] * rot[2 * 3 + j] + tr[j];
  double a = std::atan2(e[0], e[1]) * (180.0 / kPi);
  a = std::fmod(a, 360.0);
  if (a < 0) a += 360.0;
  *az = a;
  *el = std::atan2(e[2], std::hypot(e[0], e[1])) * (180.0 / kPi);
}

// Interpolated horizon at an azimuth (psr.py:127-134; lightmap.py:116-120).
inline double horizon_at(const float* hz, int N, double az) {
  double pos = az * (N / 360.0);
  double fl = std::floor(pos);
  int64_t lower = static_cast<int64_t>(fl) % N;
  if (lower < 0) lower += N;
  int64_t upper = (lower + 1) % N;
  double f = pos - fl;
  return double(hz[lower]) + f * (double(hz[upper]) - double(hz[lower]));
}

// ---------------------------------------------------------------------------
// Horizon tile codec (file_format.py:39-73 encoder, 135-182 decoder).
// ---------------------------------------------------------------------------

const float kShortScale = 32767.0f / 50.0f;
const float kElevationScale = 50.0f / 32767.0f;

int encode_pixel(const float* hz, int N, uint8_t* out) {
  int o = 0, previous = 0;
  for (int i = 0; i < N; ++i) {
    double v = double(hz[i]);
    v = std::max(-50.0, std::min(50.0, v));
    double scaled = double(float(v) * kShortScale);
    long q = static_cast<long>(scaled >= 0 ? std::floor(scaled + 0.5) : std::ceil(scaled - 0.5));
    q = std::max(-32767L, std::min(32767L, q));
    if (i == 0) {
      previous = int(q);
      out[o++] = uint8_t((q >> 8) & 0xFF);
      out[o++] = uint8_t(q & 0xFF);
      continue;
    }
    long dlt = std::max(-16384L, std::min(16383L, q - previous));
    previous = int(((previous + dlt + 32768) % 65536 + 65536) % 65536) - 32768;
    if (0 <= dlt && dlt < 63) {
      out[o++] = uint8_t(dlt);
    } else if (-64 <= dlt && dlt < 0) {
      out[o++] = uint8_t(dlt & 0x7F);
    } else {
      out[o++] = uint8_t(((dlt >> 8) & 0xFF) | 0x80);
      out[o++] = uint8_t(dlt & 0xFF);
    }
  }
  return o;
}

}  // namespace

// ===========================================================================
// C ABI of the oracle
// ===========================================================================
extern "C" {

int lsho_pyramid_layout(int32_t width, int32_t height, int32_t* levels, int32_t cap,
                        int32_t* level_count, int64_t* mip_count) {
  if (width <= 0 || height <= 0) return -1;
  int n = pyramid_layout(width, height, levels, cap, mip_count);
  if (n < 0) return -1;
  *level_count = n;
  return 0;
}

// build_max_pyramid (pyramid.py:110-138).
int lsho_pyramid_build(const float* dem, int32_t width, int32_t height, float* mips) {
  int32_t levels[64 * 4];
  int64_t total;
  int L = pyramid_layout(width, height, levels, 64, &total);
  if (L < 0) return -1;
  const float* src = dem;
  int sw = width, sh = height;
  for (int l = 1; l < L; ++l) {
    float* dst = mips + levels[l * 4 + 1];
    int dw = levels[l * 4 + 2], dh = levels[l * 4 + 3];
    for (int r = 0; r < dh; ++r)
      for (int c = 0; c < dw; ++c) {
        float m = -32000.0f;
        for (int yy = r * 4; yy < std::min(r * 4 + 4, sh); ++yy)
          for (int xx = c * 4; xx < std::min(c * 4 + 4, sw); ++xx) {
            float v = src[static_cast<int64_t>(yy) * sw + xx];
            if (valid_elevation(v) && v > m) m = v;
          }
        dst[static_cast<int64_t>(r) * dw + c] = m;
      }
    src = dst;
    sw = dw;
    sh = dh;
  }
  return 0;
}

// _build_samples for one job (geometry_numba.py:226-297). samples: float64[16][8].
int lsho_build_samples(const lsho_dem* dem, const double* observer, const double* direction,
                       double start, double maximum, double* samples, int32_t* count) {
  Geo g = make_geo(*dem);
  *count = build_samples(observer, direction, start, maximum, g, samples);
  return 0;
}

// _fit_one (geometry_numba.py:142-169).
int lsho_fit_segment(const double* samples, int32_t count, double map_resolution,
                     const double* observer, double radius, double fallback_start,
                     float* segment) {
  fit_one(samples, count, map_resolution, observer, radius, fallback_start, segment);
  return 0;
}

// build_subpatch_segments_numba (geometry.py:575-673): float32[N][C][D][18].
// exact_chain = 0 follows the production Numba path: the next DEM starts at
// float32(s_end_km) * 1000 (geometry.py:662-663). exact_chain = 1 follows the
// Python cache path and the C# captures: it starts at the last sample's float64
// distance (geometry.py:540-543).
int lsho_segments(const lsho_dem* dems, int32_t D, int32_t tile_column, int32_t tile_row,
                  int32_t tile_width, int32_t N, double max_distance, double obs_h,
                  int32_t subpatch, int32_t exact_chain, float* out, int32_t nthreads) {
  if (D <= 0 || N <= 0 || tile_width <= 0 || subpatch <= 0) return -1;
  Geo primary = make_geo(dems[0]);
  int count = tile_width / subpatch + 2;
  int C = count * count;
  std::vector<std::pair<int, int>> slot_center(C);
  std::map<std::pair<int, int>, int> unique_index;
  std::vector<std::pair<int, int>> uniques;
  std::vector<int> slot_unique(C);
  for (int i = 0; i < C; ++i) {
    int gr = i / count, gc = i % count;
    int rc = tile_column + (gc - 1) * subpatch + subpatch / 2;
    int rr = tile_row + (gr - 1) * subpatch + subpatch / 2;
    auto key = std::make_pair(clamp_center(rc, primary.w, subpatch),
                              clamp_center(rr, primary.h, subpatch));
    auto it = unique_index.find(key);
    if (it == unique_index.end()) {
      it = unique_index.emplace(key, int(uniques.size())).first;
      uniques.push_back(key);
    }
    slot_unique[i] = it->second;
  }
  int U = int(uniques.size());
  std::vector<double> observers(U * 3), radii(U), dirs(static_cast<size_t>(U) * N * 3);
  for (int u = 0; u < U; ++u)
    center_frame(primary, uniques[u].first, uniques[u].second, N, obs_h, &observers[u * 3],
                 &radii[u], &dirs[static_cast<size_t>(u) * N * 3]);
  std::vector<Geo> geos(D);
  std::vector<double> res(D), limit(D);
  for (int d = 0; d < D; ++d) {
    geos[d] = make_geo(dems[d]);
    res[d] = map_resolution(geos[d].t);
    double size_m = std::min(geos[d].w * res[d], geos[d].h * res[d]);
    limit[d] = std::min(max_distance, size_m * 1.2);
  }
  std::vector<float> unique_out(static_cast<size_t>(U) * N * D * 18, 0.0f);
  int64_t jobs = int64_t(U) * N;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t j = 0; j < jobs; ++j) {
    int u = int(j / N), a = int(j % N);
    double start = 1.0;
    double samples[kMaxSamples * kSampleFields];
    for (int d = 0; d < D; ++d) {
      if (!(start < max_distance)) break;
      int cnt = build_samples(&observers[u * 3], &dirs[(static_cast<size_t>(u) * N + a) * 3],
                              start, limit[d], geos[d], samples);
      float* seg = &unique_out[((static_cast<size_t>(u) * N + a) * D + d) * 18];
      fit_one(samples, cnt, res[d], &observers[u * 3], radii[u], start, seg);
      if (exact_chain)
        start = std::min(limit[d], cnt ? samples[(cnt - 1) * kSampleFields] : start);
      else
        start = std::min(limit[d], double(seg[13]) * 1000.0);
    }
  }
  for (int a = 0; a < N; ++a)
    for (int i = 0; i < C; ++i) {
      int u = slot_unique[i];
      std::memcpy(out + ((static_cast<size_t>(a) * C + i) * D) * 18,
                  &unique_out[((static_cast<size_t>(u) * N + a) * D) * 18],
                  sizeof(float) * D * 18);
    }
  return 0;
}

// Mode bits of the march oracle.
enum { LSHO_FLOAT32 = 1, LSHO_CAP = 2, LSHO_REFERENCE_SLOPE = 4 };

// hierarchy_maximum for one already-interpolated segment (cuda_backend.py:656-874).
// trace: float32[trace_cap][12] (nullable); returns the maximum slope in *result.
int lsho_march_ray(const float* segment, const lsho_pyramid* pyr, double observer_z,
                   double radius, double map_resolution, int32_t pass_index, double initial,
                   int32_t mode, float* trace, int32_t trace_cap, int32_t* trace_count,
                   double* result) {
  MarchDem md{pyr->dem.elevation, pyr->mips, pyr->levels, pyr->level_count, pyr->dem.width,
              pyr->dem.height};
  MarchOpts o{};
  o.cap = mode & LSHO_CAP;
  o.stable_slope = !(mode & LSHO_REFERENCE_SLOPE);
  o.trace = trace;
  o.trace_cap = trace_cap;
  if (mode & LSHO_FLOAT32) {
    float sg[18];
    for (int i = 0; i < 18; ++i) sg[i] = segment[i];
    *result = hierarchy_maximum<float>(sg, md, float(observer_z), float(radius),
                                       float(map_resolution), pass_index, float(initial), o);
  } else {
    double sg[18];
    for (int i = 0; i < 18; ++i) sg[i] = segment[i];
    *result = hierarchy_maximum<double>(sg, md, observer_z, radius, map_resolution, pass_index,
                                        initial, o);
  }
  if (trace_count) *trace_count = o.trace_count;
  return 0;
}

// One horizon tile over every DEM pass: float32 slopes [tile_h*tile_w][N] (pixel-major,
// cuda_backend.py:1018-1031; padded pixels stay -inf). counters (nullable):
// int64[3] = level-0 samples, bound tests, trace rows.
int lsho_horizon_tile(const lsho_pyramid* pyrs, int32_t D, const float* segments, int32_t N,
                      int32_t tile_column, int32_t tile_row, int32_t tile_width,
                      int32_t tile_height, int32_t subpatch, double obs_h, int32_t mode,
                      const int32_t* pixels, int32_t pixel_count, float* slopes,
                      int64_t* counters, int32_t nthreads) {
  if (D <= 0 || N <= 0) return -1;
  int count = tile_width / subpatch + 2;
  int C = count * count;
  int P = pixels ? pixel_count : tile_width * tile_height;
  int64_t total = int64_t(P) * N;
  int64_t s_samples = 0, s_tests = 0, s_steps = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : s_samples, s_tests, s_steps)
  for (int64_t j = 0; j < total; ++j) {
    int pi = int(j / N), az = int(j % N);
    int pixel = pixels ? pixels[pi] : pi;
    int row = pixel / tile_width, column = pixel % tile_width;
    float* dst = slopes + j;
    if (tile_column + column >= pyrs[0].dem.width || tile_row + row >= pyrs[0].dem.height) {
      *dst = -std::numeric_limits<float>::infinity();
      continue;
    }
    MarchOpts o{};
    o.cap = mode & LSHO_CAP;
    o.stable_slope = !(mode & LSHO_REFERENCE_SLOPE);
    if (mode & LSHO_FLOAT32) {
      *dst = ray_all_passes<float>(pyrs, D, segments, N, C, az, column, row, tile_column,
                                   tile_row, tile_width, subpatch, obs_h, o);
    } else {
      *dst = float(ray_all_passes<double>(pyrs, D, segments, N, C, az, column, row,
                                          tile_column, tile_row, tile_width, subpatch, obs_h,
                                          o));
    }
    s_samples += o.samples;
    s_tests += o.tests;
    s_steps += o.steps;
  }
  if (counters) {
    counters[0] = s_samples;
    counters[1] = s_tests;
    counters[2] = s_steps;
  }
  return 0;
}

// Float64 trace of one (pixel, azimuth) ray of a tile over every pass (rows as in
// lsho_march_ray, concatenated pass after pass). Returns the row count.
int lsho_trace_tile_ray(const lsho_pyramid* pyrs, int32_t D, const float* segments, int32_t N,
                        int32_t tile_column, int32_t tile_row, int32_t tile_width,
                        int32_t subpatch, double obs_h, int32_t pixel, int32_t az,
                        float* trace, int32_t trace_cap, double* result) {
  int count = tile_width / subpatch + 2;
  MarchOpts o{};
  o.stable_slope = true;
  o.trace = trace;
  o.trace_cap = trace_cap;
  *result = ray_all_passes<double>(pyrs, D, segments, N, count * count, az, pixel % tile_width,
                                   pixel / tile_width, tile_column, tile_row, tile_width,
                                   subpatch, obs_h, o);
  return o.trace_count;
}

// Same as lsho_horizon_tile but returns float64 slopes (method mode only), so
// parity figures are stated against the unrounded float64 result.
int lsho_horizon_tile_f64(const lsho_pyramid* pyrs, int32_t D, const float* segments,
                          int32_t N, int32_t tile_column, int32_t tile_row,
                          int32_t tile_width, int32_t tile_height, int32_t subpatch,
                          double obs_h, const int32_t* pixels, int32_t pixel_count,
                          double* slopes, int64_t* counters, int32_t nthreads) {
  if (D <= 0 || N <= 0) return -1;
  int count = tile_width / subpatch + 2;
  int C = count * count;
  int P = pixels ? pixel_count : tile_width * tile_height;
  int64_t total = int64_t(P) * N;
  int64_t s_samples = 0, s_tests = 0, s_steps = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : s_samples, s_tests, s_steps)
  for (int64_t j = 0; j < total; ++j) {
    int pi = int(j / N), az = int(j % N);
    int pixel = pixels ? pixels[pi] : pi;
    int row = pixel / tile_width, column = pixel % tile_width;
    if (tile_column + column >= pyrs[0].dem.width || tile_row + row >= pyrs[0].dem.height) {
      slopes[j] = -std::numeric_limits<double>::infinity();
      continue;
    }
    MarchOpts o{};
    o.stable_slope = true;
    slopes[j] = ray_all_passes<double>(pyrs, D, segments, N, C, az, column, row, tile_column,
                                       tile_row, tile_width, subpatch, obs_h, o);
    s_samples += o.samples;
    s_tests += o.tests;
    s_steps += o.steps;
  }
  if (counters) {
    counters[0] = s_samples;
    counters[1] = s_tests;
    counters[2] = s_steps;
  }
  return 0;
}

// The frozen-schedule evaluation of a tile (see MarchOpts): the march runs on
// `pyrs`, every level-0 sample is re-evaluated on eval_level0[pass] (float32
// grids shaped like the passes' DEMs), out[P][N] = the maximum re-evaluated
// slope (float64), -inf when the ray has no sample. This is the convention of
// the adjoint (sample positions and culls held at the forward's), used to check
// it by finite differences.
int lsho_horizon_tile_frozen(const lsho_pyramid* pyrs, const float* const* eval_level0, int32_t D,
                             const float* segments, int32_t N, int32_t tile_col,
                             int32_t tile_row, double obs_h, const int32_t* pixels, int32_t P,
                             double* out, int32_t nthreads) {
  if (!pyrs || !eval_level0 || !segments || !pixels || !out || D < 1 || N < 1 || P < 1) return -1;
  const int count = 128 / 8 + 2, C = count * count;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t job = 0; job < static_cast<int64_t>(P) * N; ++job) {
    const int k = static_cast<int>(job / N), az = static_cast<int>(job % N);
    const int column = pixels[k] % 128, row = pixels[k] / 128;
    MarchOpts o{};
    o.stable_slope = true;
    ray_all_passes<double>(pyrs, D, segments, N, C, az, column, row, tile_col, tile_row, 128, 8,
                           obs_h, o, eval_level0);
    out[job] = o.eval_current > -1e29 ? o.eval_current : -std::numeric_limits<double>::infinity();
  }
  return 0;
}

// _sun_fraction_reference (lightmap.py:32-73), float32 or float64 slices.
int lsho_sun_fraction(const float* horizon, int32_t N, double az_deg, double el_deg,
                      int32_t float32_mode, double* fraction) {
  if (float32_mode)
    *fraction = sun_fraction<float>(horizon, N, float(az_deg), float(el_deg));
  else
    *fraction = sun_fraction<double>(horizon, N, az_deg, el_deg);
  return 0;
}

// Per-pixel azimuth/elevation of body vectors in the pixel frame (psr.py:29-58).
int lsho_azimuth_elevation(const lsho_dem* dem, int32_t row, int32_t column,
                           const double* vectors, int32_t T, double* azimuth,
                           double* elevation) {
  double rot[9], tr[3];
  pixel_frame(*dem, row, column, rot, tr);
  for (int t = 0; t < T; ++t)
    azimuth_elevation(vectors + 3 * t, rot, tr, &azimuth[t], &elevation[t]);
  return 0;
}

// Lighting over one patch in float64 (lightmap.py:162-215; psr.py:94-140;
// lightmap.py:116-121). horizons: float32[128][128][N] (pixel-major tile).
// Outputs (each nullable): fraction float32[T][vh][vw] (sun fraction, float64
// slices rounded to float32); bytes uint8[T][vh][vw] = uint8(255*fraction);
// margin float32[T][vh][vw] (elevation minus interpolated horizon);
// sun_mean / visible_fraction float32[vh][vw] (mean fraction; mean of margin >= 0).
int lsho_lighting_patch(const lsho_dem* dem, const float* horizons, int32_t N,
                        const double* vectors, int32_t T, int32_t tile_y, int32_t tile_x,
                        int32_t vh, int32_t vw, int32_t float32_slices, float* fraction,
                        uint8_t* bytes, float* margin, float* sun_mean,
                        float* visible_fraction, int32_t nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < vh * vw; ++p) {
    int ly = p / vw, lx = p % vw;
    double rot[9], tr[3];
    pixel_frame(*dem, tile_y + ly, tile_x + lx, rot, tr);
    const float* hz = horizons + (static_cast<int64_t>(ly) * 128 + lx) * N;
    double acc = 0.0;
    int64_t vis = 0;
    for (int t = 0; t < T; ++t) {
      double az, el;
      azimuth_elevation(vectors + 3 * t, rot, tr, &az, &el);
      int64_t o = (static_cast<int64_t>(t) * vh + ly) * vw + lx;
      double f = float32_slices ? double(sun_fraction<float>(hz, N, float(az), float(el)))
                                : sun_fraction<double>(hz, N, az, el);
      acc += f;
      if (fraction) fraction[o] = float(f);
      if (bytes) bytes[o] = uint8_t(float(255.0f * float(f)));
      double m = el - horizon_at(hz, N, az);
      if (margin) margin[o] = float(m);
      if (m >= 0.0) ++vis;
    }
    if (sun_mean) sun_mean[p] = float(acc / T);
    if (visible_fraction) visible_fraction[p] = float(double(vis) / T);
  }
  return 0;
}

// compute_psr_patch_reference (psr.py:94-140): 255 = PSR, 0 = lit at least once.
// max_margin (nullable): float64[vh][vw] max over vectors of elevation - horizon.
int lsho_psr_patch(const lsho_dem* dem, const float* horizons, int32_t N,
                   const double* vectors, int32_t T, int32_t tile_y, int32_t tile_x,
                   int32_t vh, int32_t vw, double sun_angular_size_deg, uint8_t* out,
                   double* max_margin, int32_t nthreads) {
  double threshold = -sun_angular_size_deg / 2.0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < vh * vw; ++p) {
    int ly = p / vw, lx = p % vw;
    double rot[9], tr[3];
    pixel_frame(*dem, tile_y + ly, tile_x + lx, rot, tr);
    const float* hz = horizons + (static_cast<int64_t>(ly) * 128 + lx) * N;
    uint8_t v = 255;
    double best = -1e300;
    for (int t = 0; t < T; ++t) {
      double az, el;
      azimuth_elevation(vectors + 3 * t, rot, tr, &az, &el);
      double m = el - horizon_at(hz, N, az);
      best = std::max(best, m);
      if (m > threshold) v = 0;
    }
    out[p] = v;
    if (max_margin) max_margin[p] = best;
  }
  return 0;
}

// reduce_sun_vectors_for_psr (psr.py:61-91): indices in first-seen order.
int lsho_psr_reduce(const lsho_dem* dem, const double* vectors, int32_t T, int32_t N,
                    int64_t* indices, int32_t* count) {
  int H = dem->height, W = dem->width;
  int pts[5][2] = {{0, 0}, {0, W - 1}, {H - 1, 0}, {H - 1, W - 1}, {H / 2, W / 2}};
  std::vector<int64_t> selected(5 * static_cast<size_t>(N), -1);
  std::vector<double> maxima(5 * static_cast<size_t>(N), -std::numeric_limits<double>::infinity());
  for (int k = 0; k < 5; ++k) {
    double rot[9], tr[3];
    pixel_frame(*dem, pts[k][0], pts[k][1], rot, tr);
    for (int t = 0; t < T; ++t) {
      double az, el;
      azimuth_elevation(vectors + 3 * t, rot, tr, &az, &el);
      int64_t b = static_cast<int64_t>(az * (N / 360.0)) % N;
      if (b < 0) b += N;
      size_t idx = static_cast<size_t>(k) * N + b;
      if (el > maxima[idx]) {
        maxima[idx] = el;
        selected[idx] = t;
      }
    }
  }
  std::vector<char> seen(T, 0);
  int n = 0;
  for (int64_t v : selected)
    if (v >= 0 && !seen[v]) {
      seen[v] = 1;
      indices[n++] = v;
    }
  *count = n;
  return 0;
}

// .cbin encoder (file_format.py:39-73). out: uint8[P][2N], lengths: uint16[P].
int lsho_cbin_encode(const float* degrees, int32_t P, int32_t N, uint8_t* out,
                     uint16_t* lengths) {
  for (int p = 0; p < P; ++p)
    lengths[p] = uint16_t(encode_pixel(degrees + static_cast<int64_t>(p) * N, N,
                                       out + static_cast<int64_t>(p) * 2 * N));
  return 0;
}

// .cbin decoder (file_format.py:135-182) with its structural validation.
// Returns 0, or -1 on any structural error.
int lsho_cbin_decode(const uint8_t* payload, int64_t size, int32_t P, int32_t N, float* out) {
  int64_t off = 0;
  for (int p = 0; p < P; ++p) {
    if (off + 2 > size) return -1;
    int length = payload[off] | (payload[off + 1] << 8);
    off += 2;
    if (length < 2 || length > 2 * N) return -1;
    int64_t end = off + length;
    if (end > size) return -1;
    int acc = (payload[off] << 8) | payload[off + 1];
    if (acc >= 32768) acc -= 65536;
    off += 2;
    float* o = out + static_cast<int64_t>(p) * N;
    o[0] = float(acc) * kElevationScale;
    int i = 1;
    while (off < end && i < N) {
      int first = payload[off++];
      int dl;
      if (first & 0x80) {
        if (off >= end) return -1;
        int high = ((first << 1) & 0x80) | (first & 0x7F);
        dl = (high << 8) | payload[off++];
        if (dl >= 32768) dl -= 65536;
      } else {
        dl = first & 0x7F;
        if (dl & 0x40) dl -= 0x80;
      }
      acc = ((acc + dl + 32768) % 65536 + 65536) % 65536 - 32768;
      o[i++] = float(acc) * kElevationScale;
    }
    if (i != N || off != end) return -1;
  }
  return off == size ? 0 : -1;
}

}  // extern "C"
