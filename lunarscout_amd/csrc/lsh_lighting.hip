// lsh_lighting.hip -- lighting from stored horizons: sun fraction (finite disk),
// Earth margin / visibility and the PSR mask, in one streamed pass per pixel.
//
// The per-pixel arithmetic is lsh_lighting_core.h (shared with the host path).
// One thread walks one series (Sun, Earth or PSR) of one pixel: the grid's z
// index picks the series, so a launch holds up to three independent work items
// per pixel and the Sun walks -- the longest -- are dispatched first. All
// threads of a workgroup read the same Moon-ME vector (a uniform, scalar load)
// and each reads only its own horizon row. The horizons may be pixel-major
// [pixel][azimuth] (the tile file layout) or azimuth-major [azimuth][pixel], in
// which case the 64 lanes of a wavefront read one contiguous 256-byte run per
// tap.

#include <thread>
#include <vector>

#include "lsh_lighting_core.h"

namespace lsh {
namespace {

// Wavefronts per SIMD the lighting kernel's registers are sized for. At 4 the
// walk keeps its state in 113 VGPRs without spilling; 5 and 6 spill 56-84 B per
// lane and measured 10-18 % slower on a year of lighting over 24 tiles.
#ifndef LSH_LIGHT_WAVES_PER_EU
#define LSH_LIGHT_WAVES_PER_EU 4
#endif

// The series each z slice of the grid walks, in launch order.
struct SeriesList {
  unsigned series[3];
};

template <bool kAzimuthMajor>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LSH_LIGHT_WAVES_PER_EU)))
lighting_kernel(LightArgs a, SeriesList list) {
  // each thread's sector tables, [sector][thread] (consecutive lanes, consecutive banks)
  __shared__ float s_dark[kSectors][256], s_lit[kSectors][256];
  const int tile = blockIdx.y;
  const int pixel = blockIdx.x * blockDim.x + threadIdx.x;
  if (pixel >= kPatchPixels) return;
  const SectorTable sectors{&s_dark[0][threadIdx.x], &s_lit[0][threadIdx.x], 256};
  light_pixel<kAzimuthMajor>(a, tile, pixel, a.tiles[tile], sectors, list.series[blockIdx.z]);
}

}  // namespace

hipError_t launch_lighting(const LightArgs& a, hipStream_t stream) {
  SeriesList list{{0u, 0u, 0u}};
  unsigned count = 0;
  // a series is walked where one of its outputs is requested
  if (a.sun_bytes || a.sun_fraction || a.sun_mean) list.series[count++] = kSunSeries;
  if (a.earth_margin || a.earth_visible) list.series[count++] = kEarthSeries;
  if (a.psr || a.psr_margin) list.series[count++] = kPsrSeries;
  if (count == 0) return hipSuccess;
  dim3 block(256), grid(kPatchPixels / 256, a.tile_count, count);
  if (a.azimuth_major)
    hipLaunchKernelGGL(lighting_kernel<true>, grid, block, 0, stream, a, list);
  else
    hipLaunchKernelGGL(lighting_kernel<false>, grid, block, 0, stream, a, list);
  return hipGetLastError();
}

}  // namespace lsh

namespace lsh {

// The host path: the same per-pixel arithmetic over (tile, pixel) ranges on
// `threads` host threads (the "cpu" backend of the lighting products).
void run_lighting_host(const LightArgs& a, int threads) {
  const int64_t items = static_cast<int64_t>(a.tile_count) * kPatchPixels;
  int n = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
  n = n < 1 ? 1 : (n > 256 ? 256 : n);
  auto body = [&a, items, n](int t) {
    const int64_t begin = items * t / n, end = items * (t + 1) / n;
    float dark[kSectors], lit[kSectors];
    const SectorTable sectors{dark, lit, 1};
    for (int64_t i = begin; i < end; ++i) {
      const int tile = static_cast<int>(i / kPatchPixels), pixel = static_cast<int>(i % kPatchPixels);
      if (a.azimuth_major)
        light_pixel<true>(a, tile, pixel, a.tiles[tile], sectors);
      else
        light_pixel<false>(a, tile, pixel, a.tiles[tile], sectors);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < n; ++t) pool.emplace_back(body, t);
  body(0);
  for (auto& th : pool) th.join();
}

}  // namespace lsh
