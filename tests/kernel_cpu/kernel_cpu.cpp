// CPU build of the HIP march core (lunarscout_amd/csrc/lsh_march_core.h).
//
// Test infrastructure: runs the GPU kernel's per-ray arithmetic on the host so
// the tests can (1) state the float32 kernel's parity against the float64
// oracle without a GPU and (2) check the GPU output bit for bit. It follows the
// per-pixel driver of march_kernel in lsh_march.hip.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "../../lunarscout_amd/csrc/lsh_march_core.h"
#include "../../lunarscout_amd/csrc/lsh_math.h"

namespace {
struct PassTables {
  int32_t off[4][16], lw[4][16], lh[4][16];
  float colres[4], mapres[4];
  int32_t all_valid[4];
  int32_t bound_offset[4];
  // level 0, the mips, then the 2x2 bound of every level, as the device holds them
  std::vector<float> combined[4];
};

bool valid(float v) { return std::isfinite(v) && v > -20000.0f; }

int res_exponent(float res) {
  int e = 0;
  (void)std::frexp(double(res), &e);
  return e;
}

double res_mantissa(float res) {
  int e = 0;
  return std::frexp(double(res), &e);
}
}  // namespace

extern "C" {

typedef struct {
  const float* elevation;
  int32_t width;
  int32_t height;
  double geo_transform[6];
  double projection[6];
} kc_dem;

typedef struct {
  kc_dem dem;
  const float* mips;
  const int32_t* levels;  // [L][4]
  int32_t level_count;
} kc_pyramid;

}  // extern "C"

namespace {

// One ray over every pass; S = schedule precision, V = value precision.
template <class S, class V, class Tracer>
float ray(const kc_pyramid* pyrs, int D, const float* segments, int C, int az, int column,
          int row, int tile_column, int tile_row, int tile_width, int subpatch, float obs_h,
          bool cap, bool guarded, const PassTables& t, int32_t* best_out,
          lsh::MarchCounters& cnt, const Tracer& tr) {
  const kc_dem& primary = pyrs[0].dem;
  lsh::PixelSelection sel = lsh::lsh_select(column, row, tile_column, tile_row, tile_width,
                                            subpatch, primary.width, primary.height);
  V oz = lsh::lsh_observer_terrain<V>(primary.elevation, primary.width, primary.height,
                                      tile_column + column, tile_row + row) + V(obs_h);
  V current = V(-1e30);
  int32_t best = -1;
  int64_t stride = int64_t(D) * 18;
  const float* seg_az = segments + int64_t(az) * C * stride;
  for (int p = 0; p < D; ++p) {
    const float* base = seg_az + p * 18;
    lsh::PassSegment<S> g = lsh::lsh_interpolate<S>(
        base + sel.i00 * stride, base + sel.i10 * stride, base + sel.i01 * stride,
        base + sel.i11 * stride, sel, column, row, t.colres[0] / t.colres[p]);
    lsh::MarchView v{t.combined[p].data(), uint32_t(4 * t.combined[p].size()), t.bound_offset[p],
                     t.off[p],
                     t.lw[p], t.lh[p], pyrs[p].dem.width, pyrs[p].dem.height,
                     pyrs[p].level_count, t.mapres[p], float(pyrs[p].dem.projection[0]),
                     t.all_valid[p], double(t.mapres[p]), res_exponent(t.mapres[p]),
                     res_mantissa(t.mapres[p]), 0u, 0};
    if (guarded && pyrs[p].level_count > 2) {  // the level-2 bounds, as march_dem sets them
      v.box_offset = uint32_t(t.bound_offset[p] + t.off[p][2]);
      v.box_width = t.lw[p][2];
    }
    if (cap)
      lsh::lsh_march_pass<S, V, true, true>(g, v, oz, p, current, best, cnt, tr);
    else
      lsh::lsh_march_pass<S, V, false, true>(g, v, oz, p, current, best, cnt, tr);
  }
  *best_out = best;
  return current > V(-1e29) ? float(current) : -std::numeric_limits<float>::infinity();
}

PassTables tables(const kc_pyramid* pyrs, int D) {
  PassTables t;
  for (int p = 0; p < D; ++p) {
    const int64_t level0 = int64_t(pyrs[p].dem.width) * pyrs[p].dem.height;
    int64_t mip_count = 0;
    for (int l = 1; l < pyrs[p].level_count; ++l)
      mip_count += int64_t(pyrs[p].levels[l * 4 + 2]) * pyrs[p].levels[l * 4 + 3];
    t.combined[p].assign(pyrs[p].dem.elevation, pyrs[p].dem.elevation + level0);
    t.combined[p].insert(t.combined[p].end(), pyrs[p].mips, pyrs[p].mips + mip_count);
    if (mip_count == 0) t.combined[p].push_back(0.0f);  // the device's one-float placeholder
    for (int l = 0; l < 16; ++l) {
      bool used = l < pyrs[p].level_count;
      t.off[p][l] = used && l > 0 ? int32_t(level0 + pyrs[p].levels[l * 4 + 1]) : 0;
      t.lw[p][l] = used ? pyrs[p].levels[l * 4 + 2] : 1;
      t.lh[p][l] = used ? pyrs[p].levels[l * 4 + 3] : 1;
    }
    // bound of each cell: max of the valid values of its clipped 2x2 block, else -32000
    const int64_t n = int64_t(t.combined[p].size());
    t.bound_offset[p] = int32_t(n);
    t.combined[p].resize(2 * n, 0.0f);
    for (int l = 0; l < pyrs[p].level_count; ++l) {
      const float* src = t.combined[p].data() + t.off[p][l];
      float* dst = t.combined[p].data() + n + t.off[p][l];
      const int w = t.lw[p][l], h = t.lh[p][l];
      for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
          const int x1 = x + 1 < w ? x + 1 : w - 1, y1 = y + 1 < h ? y + 1 : h - 1;
          const float taps[4] = {src[int64_t(y) * w + x], src[int64_t(y) * w + x1],
                                 src[int64_t(y1) * w + x], src[int64_t(y1) * w + x1]};
          float m = -32000.0f;
          for (float v : taps)
            if (valid(v) && v > m) m = v;
          dst[int64_t(y) * w + x] = m;
        }
    }
    const double* g = pyrs[p].dem.geo_transform;
    float t1 = float(g[1]), t4 = float(g[4]), t2 = float(g[2]), t5 = float(g[5]);
    t.colres[p] = std::sqrt(t1 * t1 + t4 * t4);
    t.mapres[p] = (t.colres[p] + std::sqrt(t2 * t2 + t5 * t5)) * 0.5f;
    t.all_valid[p] = 1;
    const float* e = pyrs[p].dem.elevation;
    for (int64_t i = 0, n = int64_t(pyrs[p].dem.width) * pyrs[p].dem.height; i < n; ++i)
      if (!(std::isfinite(e[i]) && e[i] > -20000.0f)) {
        t.all_valid[p] = 0;
        break;
      }
  }
  return t;
}

template <class Tracer>
float ray_mode(int mode, const kc_pyramid* pyrs, int D, const float* segments, int C, int az,
               int column, int row, int tile_column, int tile_row, int tile_width, int subpatch,
               float obs_h, const PassTables& t, int32_t* best, lsh::MarchCounters& cnt,
               const Tracer& tr) {
  bool cap = mode & 1, guarded = !(mode & 8);
  switch (mode & 6) {
    case 0:
      return ray<float, float>(pyrs, D, segments, C, az, column, row, tile_column, tile_row,
                               tile_width, subpatch, obs_h, cap, guarded, t, best, cnt, tr);
    case 2:
      return ray<double, float>(pyrs, D, segments, C, az, column, row, tile_column, tile_row,
                                tile_width, subpatch, obs_h, cap, guarded, t, best, cnt, tr);
    default:
      return ray<double, double>(pyrs, D, segments, C, az, column, row, tile_column, tile_row,
                                 tile_width, subpatch, obs_h, cap, guarded, t, best, cnt, tr);
  }
}

}  // namespace

extern "C" {

// flags: 1 = cap, 2 = float64 schedule, 4 = float64 values, 8 = no guarded skips.
// counters: int64[3] (samples, tests, steps).
int kcpu_horizon_tile(const kc_pyramid* pyrs, int32_t D, const float* segments, int32_t N,
                      int32_t tile_column, int32_t tile_row, int32_t tile_width,
                      int32_t tile_height, int32_t subpatch, float obs_h, int32_t flags,
                      const int32_t* pixels, int32_t pixel_count, float* out, int32_t* winners,
                      int64_t* counters, int32_t nthreads) {
  if (D < 1 || D > 4) return -1;
  int count = tile_width / subpatch + 2;
  int C = count * count;
  int P = pixels ? pixel_count : tile_width * tile_height;
  PassTables t = tables(pyrs, D);
  const kc_dem& primary = pyrs[0].dem;
  int64_t total = int64_t(P) * N;
  int64_t cs = 0, ct = 0, cst = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : cs, ct, cst)
  for (int64_t j = 0; j < total; ++j) {
    int pi = int(j / N), az = int(j % N);
    int pixel = pixels ? pixels[pi] : pi;
    int row = pixel / tile_width, column = pixel % tile_width;
    if (tile_column + column >= primary.width || tile_row + row >= primary.height) {
      out[j] = -std::numeric_limits<float>::infinity();
      if (winners) winners[j] = -1;
      continue;
    }
    lsh::MarchCounters cnt{0, 0, 0};
    int32_t best = -1;
    out[j] = ray_mode(flags, pyrs, D, segments, C, az, column, row, tile_column, tile_row,
                      tile_width, subpatch, obs_h, t, &best, cnt, lsh::NoTrace());
    if (winners) winners[j] = best;
    cs += cnt.samples;
    ct += cnt.tests;
    cst += cnt.steps;
  }
  if (counters) {
    counters[0] = cs;
    counters[1] = ct;
    counters[2] = cst;
  }
  return 0;
}

struct RowTrace : lsh::NoTrace {
  float* rows;
  int cap;
  int* count;
  void row(float s, float t, int l, int cx, int cy, float px, float py, float mh, float hd,
           float sl, float adv, int act) const {
    if (*count < cap) {
      float* r = rows + *count * 12;
      r[0] = s; r[1] = t; r[2] = float(l); r[3] = float(cx); r[4] = float(cy); r[5] = px;
      r[6] = py; r[7] = mh; r[8] = hd; r[9] = sl; r[10] = adv; r[11] = float(act);
    }
    ++*count;
  }
};

// Trace of one (pixel, azimuth) ray of a tile over every pass.
int kcpu_trace_ray(const kc_pyramid* pyrs, int32_t D, const float* segments, int32_t N,
                   int32_t tile_column, int32_t tile_row, int32_t tile_width, int32_t subpatch,
                   float obs_h, int32_t flags, int32_t pixel, int32_t az, float* rows,
                   int32_t cap, float* result) {
  int count = tile_width / subpatch + 2;
  PassTables t = tables(pyrs, D);
  lsh::MarchCounters cnt{0, 0, 0};
  int n = 0;
  RowTrace tr{{}, rows, cap, &n};
  int32_t best = -1;
  *result = ray_mode(flags, pyrs, D, segments, count * count, az, pixel % tile_width,
                     pixel / tile_width, tile_column, tile_row, tile_width, subpatch, obs_h, t,
                     &best, cnt, tr);
  return n;
}

// The lighting path's float64 arctangent over n (y, x) pairs.
void kcpu_atan2(const double* y, const double* x, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = lsh::atan2_finite(y[i], x[i]);
}

// hierarchy_maximum over one raw 18-field segment (no subpatch blend), the
// form oracle.march_ray takes: the kernel core's pass with the segment used as
// all four blend corners at zero weight (exact), so constructed rays can be
// compared with the oracle trace row for trace row. flags: 1 = cap,
// 2 = float64 schedule, 4 = float64 values. Returns the row count.
int kcpu_march_segment(const float* segment, const kc_pyramid* pyr, double observer_z,
                       int32_t pass_index, double initial, int32_t flags, float* rows,
                       int32_t cap, double* result) {
  PassTables t = tables(pyr, 1);
  lsh::PixelSelection sel{0, 0, 0, 0, 0.0f, 0.0f, 0, 0, 0, 0};
  lsh::MarchView v{t.combined[0].data(), uint32_t(4 * t.combined[0].size()), t.bound_offset[0],
                   t.off[0], t.lw[0], t.lh[0], pyr->dem.width, pyr->dem.height,
                   pyr->level_count, t.mapres[0], float(pyr->dem.projection[0]),
                   t.all_valid[0], double(t.mapres[0]), res_exponent(t.mapres[0]),
                   res_mantissa(t.mapres[0]), 0u, 0};
  lsh::MarchCounters cnt{0, 0, 0};
  int n = 0;
  RowTrace tr{{}, rows, cap, &n};
  int32_t best = -1;
  auto run = [&](auto s_tag, auto v_tag) {
    using S = decltype(s_tag);
    using V = decltype(v_tag);
    lsh::PassSegment<S> g =
        lsh::lsh_interpolate<S>(segment, segment, segment, segment, sel, 0, 0, 1.0f);
    V current = __builtin_isinf(initial) ? V(-1e30) : V(initial);
    if (flags & 1)
      lsh::lsh_march_pass<S, V, true, true>(g, v, V(observer_z), pass_index, current, best, cnt,
                                            tr);
    else
      lsh::lsh_march_pass<S, V, false, true>(g, v, V(observer_z), pass_index, current, best,
                                             cnt, tr);
    *result = current > V(-1e29) ? double(current) : -std::numeric_limits<double>::infinity();
  };
  switch (flags & 6) {
    case 0: run(0.0f, 0.0f); break;
    case 2: run(0.0, 0.0f); break;
    default: run(0.0, 0.0); break;
  }
  return n;
}

}  // extern "C"
