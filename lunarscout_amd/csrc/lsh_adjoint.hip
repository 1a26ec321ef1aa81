// lsh_adjoint.hip -- DEM sensitivity of illumination: the smoothed sun-fraction
// loss, its gradient with respect to the horizons, and the adjoint of the march.
//
// The loss over tiles of pixels p and Sun vectors t is
//   L = sum_p sum_t F(p, t),   F = sum_i c_s(el + c_i - hz_i, 2 c_i) / max_photons,
// the reference's 16-slice finite-disk fraction (lightmap.py:32-73) with the
// slice clamp min(max(x, 0), 2 c_i) replaced by c_s, a C1 clamp whose two
// corners are rounded over +-s degrees (quadratic blends): c_s equals the clamp
// outside the corners, so d(fraction)/d(horizon) is exact zero where a slice is
// fully lit or fully hidden by more than s. The horizon a slice sees is the
// linear interpolation of two 0.25-degree bins, as in the reference.
//
// The adjoint of the march keeps no history: the forward stores, per (pixel,
// azimuth), only the int32 winning-sample index (pass << 24 | level-0 sample
// index). The adjoint re-marches that ray with the identical schedule up to the
// winning sample (earlier passes run whole, for their running maximum, which
// steers the culls), then differentiates the winning slope analytically with
// respect to the four bilinear DEM nodes of the sample and the observer
// terrain nodes of the pixel. Rays whose horizon bin has zero loss gradient are
// skipped. Contributions are scattered as 64-bit fixed point (value * 2^k)
// with integer atomics: integer addition is associative, so the gradient is
// bit-identical from run to run whatever order the rays finish in.
//
// The ray geometry (segments) is held fixed: the dependence of the fitted
// segments on the terrain at subpatch centres (the observer radius of the
// chord samples, a relative effect of order 1e-9) is not differentiated.

#include <cmath>

#include "lsh_internal.h"
#include "lsh_lighting_core.h"
#include "lsh_march_core.h"

namespace lsh {
namespace {

// c_s (see above) and its derivative; W = 2 c_i >= 2 s.
__device__ inline void smooth_clamp(double x, double W, double s, double& value, double& slope) {
  if (x <= -s) {
    value = 0.0;
    slope = 0.0;
  } else if (x < s) {
    value = (x + s) * (x + s) / (4.0 * s);
    slope = (x + s) / (2.0 * s);
  } else if (x <= W - s) {
    value = x;
    slope = 1.0;
  } else if (x < W + s) {
    const double r = W + s - x;
    value = W - r * r / (4.0 * s);
    slope = r / (2.0 * s);
  } else {
    value = W;
    slope = 0.0;
  }
}

struct LossArgs {
  LightArgs light;       // DEM frame, tiles, horizons [tile][16384][N] (degrees), sun vectors
  double smoothing;      // s, degrees
  double* loss;          // nullable [tile][16384]: sum_t F(p, t)
  float* grad;           // nullable [tile][16384][N]: sum_t dF/dhz(bin)
};

// One pixel: sum over times of the smoothed fraction and, optionally, its
// gradient with respect to the pixel's horizon bins (the pixel's own row, so no
// two threads write the same element).
__global__ void __launch_bounds__(256) sun_loss_kernel(LossArgs a) {
  const LightArgs& L = a.light;
  const int tile = blockIdx.y;
  const int pixel = blockIdx.x * blockDim.x + threadIdx.x;
  if (pixel >= kPatchPixels) return;
  const lsh_tile origin = L.tiles[tile];
  const int row = origin.tile_row + pixel / 128, column = origin.tile_column + pixel % 128;
  const int N = L.N;
  const int64_t row_base = (static_cast<int64_t>(tile) * kPatchPixels + pixel) * N;
  float* grad = a.grad ? a.grad + row_base : nullptr;
  if (grad)
    for (int b = 0; b < N; ++b) grad[b] = 0.0f;
  if (row >= L.dem_height || column >= L.dem_width) {
    if (a.loss) a.loss[static_cast<int64_t>(tile) * kPatchPixels + pixel] = 0.0;
    return;
  }
  const float* hz = L.horizons + row_base;
  const Frame f = pixel_frame(L, row, column);
  const double s = a.smoothing;
  double total = 0.0;
  for (int t = 0; t < L.sun_count; ++t) {
    double az, el;
    azimuth_elevation(f, L.sun + 3 * t, az, el);
    // the slice walk of sun_fraction (lightmap.py:32-73), float32 positions
    const float el_f = static_cast<float>(el);
    const DiskWalk w = disk_walk(N);
    float left_pos = (static_cast<float>(az) - L.sun_half_angle - w.half_bucket) * w.per_degree;
    int left = static_cast<int>(left_pos);
    float frac = left_pos - static_cast<float>(left);
    if (left < 0) left += N;
    else if (left >= N) left -= N;
    int right = left + 1;
    if (right >= N) right = 0;
    int lo_bin = left;
    double photons = 0.0;
    // gradient of this time, kept per bin the slices touch: the walk spans
    // frac + 15 * step < 3.03 bins (step <= 0.135 for N <= 1440), so at most 5 bins
    double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int bins[5] = {left, right, 0, 0, 0};
    int used = 2;
    int slot_lo = 0, slot_hi = 1;
    for (int i = 0; i < 16; ++i) {
      const double c = L.half_circle[i];
      // a bin without terrain (-inf) counts as -90 degrees: fully lit, zero gradient
      const double h_lo = fmax(static_cast<double>(hz[lo_bin]), -90.0);
      const double h_hi = fmax(static_cast<double>(hz[right]), -90.0);
      const double horizon = static_cast<double>(frac) * (h_hi - h_lo) + h_lo;
      double value, slope;
      smooth_clamp(static_cast<double>(el_f) + c - horizon, 2.0 * c, s, value, slope);
      photons += value;
      if (slope != 0.0) {  // d(value)/d(horizon) = -slope
        g[slot_lo] -= slope * (1.0 - static_cast<double>(frac));
        g[slot_hi] -= slope * static_cast<double>(frac);
      }
      frac = frac + w.step;
      if (frac >= 1.0f) {
        lo_bin = right;
        right += 1;
        if (right >= N) right = 0;
        frac = frac - 1.0f;
        slot_lo = slot_hi;
        bins[used] = right;
        slot_hi = used++;
      }
    }
    total += photons / L.max_photons;
    if (grad)
      for (int k = 0; k < used; ++k)
        if (g[k] != 0.0) grad[bins[k]] += static_cast<float>(g[k] / L.max_photons);
  }
  if (a.loss) a.loss[static_cast<int64_t>(tile) * kPatchPixels + pixel] = total;
}

// Captures the k-th level-0 sample of a pass and stops the march there.
struct SampleProbe : NoTrace {
  int target;
  mutable bool found = false;
  mutable int ix = 0, iy = 0;
  mutable double tx = 0.0, ty = 0.0, hd = 0.0, true_m = 0.0;
  mutable bool near = false;
  __device__ bool sample(int index, int cx, int cy, double fx, double fy, double h, double tm,
                         bool n) const {
    if (index != target) return false;
    found = true;
    ix = cx;
    iy = cy;
    tx = fx;
    ty = fy;
    hd = h;
    true_m = tm;
    near = n;
    return true;
  }
};

struct AdjointArgs {
  MarchArgs march;                        // tile, segments, DEMs (winners = forward winners)
  const float* grad_horizon;              // [16384][N] dL/d(horizon degrees)
  unsigned long long* grad_fixed[kMaxDems];  // per DEM [height][width], fixed point
  double fixed_scale;                     // 2^k
  int32_t* sample_cells;                  // nullable [16384][N]: pass << 28 | iy << 14 | ix
};

__device__ inline void scatter(unsigned long long* grid, int64_t index, double value, double scale) {
  const long long fixed = llrint(value * scale);
  if (fixed != 0) atomicAdd(grid + index, static_cast<unsigned long long>(fixed));
}

template <bool kUnused>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
adjoint_kernel(AdjointArgs a) {
  const MarchArgs& m = a.march;
  __shared__ int32_t s_off[kMaxDems][kMaxLevels];
  __shared__ int32_t s_w[kMaxDems][kMaxLevels];
  __shared__ int32_t s_h[kMaxDems][kMaxLevels];
  for (int i = threadIdx.x; i < kMaxDems * kMaxLevels; i += blockDim.x) {
    int dd = i / kMaxLevels, l = i % kMaxLevels;
    s_off[dd][l] = m.dem[dd].level_offset[l];
    s_w[dd][l] = m.dem[dd].level_width[l];
    s_h[dd][l] = m.dem[dd].level_height[l];
  }
  __syncthreads();
  // the forward kernel's mapping: 8x8 pixels x 4 azimuths per workgroup
  const int blocks = 16 * 16;
  const int group = blockIdx.x / blocks, pblock = blockIdx.x % blocks;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int az = group * 4 + wave;
  const int column = (pblock % 16) * 8 + lane % 8, row = (pblock / 16) * 8 + lane / 8;
  if (az >= m.N) return;
  const int pixel = row * 128 + column;
  const int64_t index = static_cast<int64_t>(pixel) * m.N + az;
  if (a.sample_cells) a.sample_cells[index] = -1;
  const int32_t winner = m.winners[index];
  const double dtheta = a.grad_horizon[index];
  if (winner < 0 || dtheta == 0.0) return;
  const MarchDem& primary = m.dem[0];
  const int win_pass = winner >> 24, win_index = winner & 0xFFFFFF;
  const PixelSelection sel = lsh_select(column, row, m.tile_column, m.tile_row, 128, m.subpatch,
                                        primary.width, primary.height);
  const double oz = lsh_observer_terrain<double>(primary.level0, primary.width, primary.height,
                                                 m.tile_column + column, m.tile_row + row) +
                    static_cast<double>(m.observer_elevation);
  double current = -1e30;
  int32_t best = -1;
  MarchCounters cnt{0, 0, 0};
  const int64_t seg_stride = static_cast<int64_t>(m.D) * kSegmentFields;
  const float* seg_az = m.segments + static_cast<int64_t>(az) * m.C * seg_stride;
  SampleProbe probe;
  probe.target = win_index;
  for (int pass = 0; pass <= win_pass; ++pass) {
    const MarchDem& d = m.dem[pass];
    const float* p = seg_az + pass * kSegmentFields;
    const PassSegment<double> g = lsh_interpolate<double>(
        p + sel.i00 * seg_stride, p + sel.i10 * seg_stride, p + sel.i01 * seg_stride,
        p + sel.i11 * seg_stride, sel, column, row,
        primary.column_resolution / d.column_resolution);
    MarchView view{d.level0, d.buffer_bytes, d.bound_offset, s_off[pass], s_w[pass], s_h[pass],
                   d.width, d.height, d.level_count, d.map_resolution, d.radius, d.all_valid,
                   d.resolution, d.resolution_exponent, d.resolution_mantissa,
                   d.box_offset, d.box_width};
    if (pass < win_pass)
      lsh_march_pass<double, double, false, true>(g, view, oz, pass, current, best, cnt);
    else
      lsh_march_pass<double, double, false, true, SampleProbe>(g, view, oz, pass, current, best,
                                                               cnt, probe);
  }
  if (!probe.found) return;
  const MarchDem& d = m.dem[win_pass];
  // d slope / d h and d slope / d oz of the winning sample (see the header)
  const double R = static_cast<double>(d.radius);
  double ds_dh, ds_doz, slope;
  if (probe.near) {
    ds_dh = 1.0 / probe.true_m;
    ds_doz = -ds_dh;
    slope = probe.hd / probe.true_m;
  } else {
    const double u = R + oz, tm2 = probe.true_m * probe.true_m;
    const double lz = (probe.hd * (2.0 * R + 2.0 * oz + probe.hd) - tm2) / (2.0 * u);
    const double lxx = tm2 - lz * lz;
    if (!(lxx > 0.0)) return;
    const double ds_dlz = tm2 / (lxx * sqrt(lxx));
    ds_dh = ds_dlz * (u + probe.hd) / u;
    ds_doz = ds_dlz * (-1.0 - lz / u);
    slope = lz / sqrt(lxx);
  }
  const double dtheta_ds = (180.0 / 3.14159265358979323846) / (1.0 + slope * slope);
  const double gh = dtheta * dtheta_ds * ds_dh, go = dtheta * dtheta_ds * ds_doz;
  // the sample's 2x2 nodes, clipped like the forward taps
  const int cx = probe.ix, cy = probe.iy;
  const int cx1 = cx + 1 < d.width ? cx + 1 : d.width - 1;
  const int cy1 = cy + 1 < d.height ? cy + 1 : d.height - 1;
  const double tx = probe.tx, ty = probe.ty;
  unsigned long long* grid = a.grad_fixed[win_pass];
  scatter(grid, static_cast<int64_t>(cy) * d.width + cx, gh * (1.0 - tx) * (1.0 - ty), a.fixed_scale);
  scatter(grid, static_cast<int64_t>(cy) * d.width + cx1, gh * tx * (1.0 - ty), a.fixed_scale);
  scatter(grid, static_cast<int64_t>(cy1) * d.width + cx, gh * (1.0 - tx) * ty, a.fixed_scale);
  scatter(grid, static_cast<int64_t>(cy1) * d.width + cx1, gh * tx * ty, a.fixed_scale);
  // observer terrain: the bilinear of lsh_observer_terrain at the pixel
  {
    const int w = primary.width, h = primary.height;
    double oc = fmin(fmax(static_cast<double>(m.tile_column + column), 0.0), w - 1.0001);
    double orow = fmin(fmax(static_cast<double>(m.tile_row + row), 0.0), h - 1.0001);
    const int x0 = static_cast<int>(floor(oc)), y0 = static_cast<int>(floor(orow));
    const int x1 = x0 + 1 < w ? x0 + 1 : w - 1, y1 = y0 + 1 < h ? y0 + 1 : h - 1;
    const double fx = oc - x0, fy = orow - y0;
    unsigned long long* pg = a.grad_fixed[0];
    scatter(pg, static_cast<int64_t>(y0) * w + x0, go * (1.0 - fx) * (1.0 - fy), a.fixed_scale);
    scatter(pg, static_cast<int64_t>(y0) * w + x1, go * fx * (1.0 - fy), a.fixed_scale);
    scatter(pg, static_cast<int64_t>(y1) * w + x0, go * (1.0 - fx) * fy, a.fixed_scale);
    scatter(pg, static_cast<int64_t>(y1) * w + x1, go * fx * fy, a.fixed_scale);
  }
  if (a.sample_cells) a.sample_cells[index] = (win_pass << 28) | (cy << 14) | cx;
}

}  // namespace

hipError_t launch_sun_loss(const LightArgs& light, double smoothing, double* loss, float* grad,
                           hipStream_t stream) {
  LossArgs a{light, smoothing, loss, grad};
  hipLaunchKernelGGL(sun_loss_kernel, dim3(kPatchPixels / 256, light.tile_count), dim3(256), 0,
                     stream, a);
  return hipGetLastError();
}

hipError_t launch_adjoint(const MarchArgs& march, const float* grad_horizon,
                          unsigned long long* const* grad_fixed, double fixed_scale,
                          int32_t* sample_cells, hipStream_t stream) {
  AdjointArgs a{};
  a.march = march;
  a.grad_horizon = grad_horizon;
  for (int d = 0; d < kMaxDems; ++d) a.grad_fixed[d] = d < march.D ? grad_fixed[d] : nullptr;
  a.fixed_scale = fixed_scale;
  a.sample_cells = sample_cells;
  const int groups = (march.N + 3) / 4;
  hipLaunchKernelGGL(adjoint_kernel<true>, dim3(16 * 16 * groups), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace lsh
