// lsh_pyramid.hip -- factor-four maximum pyramid of a DEM.
//
// Restates pyramid_downsample_kernel (cuda_backend.py:134-153) and
// build_max_pyramid (pyramid.py:110-138): each destination cell is the maximum of
// the valid elevations (finite and above -20000 m) of its 4x4 source block,
// -32000 when none is valid. Maximum is exact, so the result is bit-identical to
// the reference. One launch per level.

#include "lsh_internal.h"

namespace lsh {
namespace {

__device__ inline bool valid(float v) { return isfinite(v) && v > -20000.0f; }

__global__ void __launch_bounds__(256) pyramid_kernel(const float* __restrict__ src, int sw,
                                                      int sh, float* __restrict__ dst, int dw,
                                                      int dh) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  if (c >= dw || r >= dh) return;
  float m = -32000.0f;
  int x0 = c * 4, y0 = r * 4;
  for (int dy = 0; dy < 4; ++dy) {
    int y = y0 + dy;
    if (y >= sh) break;
    const float* row = src + static_cast<int64_t>(y) * sw;
    for (int dx = 0; dx < 4; ++dx) {
      int x = x0 + dx;
      if (x >= sw) break;
      float v = row[x];
      if (valid(v) && v > m) m = v;
    }
  }
  dst[static_cast<int64_t>(r) * dw + c] = m;
}

// The march's bound of a cell: the maximum of the valid values of cells
// (x..x+1, y..y+1), neighbours clipped to the level (hierarchy_bilinear_bound,
// cuda_backend.py:47-60; a clipped neighbour repeats the cell), -32000 when none
// is valid. Stored per cell so that a bound test is one load.
__global__ void __launch_bounds__(256) bound_kernel(const float* __restrict__ src, int w, int h,
                                                    float* __restrict__ dst) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= w || y >= h) return;
  int x1 = x + 1 < w ? x + 1 : w - 1, y1 = y + 1 < h ? y + 1 : h - 1;
  const float* r0 = src + static_cast<int64_t>(y) * w;
  const float* r1 = src + static_cast<int64_t>(y1) * w;
  float taps[4] = {r0[x], r0[x1], r1[x], r1[x1]};
  float m = -32000.0f;
  for (float v : taps)
    if (valid(v) && v > m) m = v;
  dst[static_cast<int64_t>(y) * w + x] = m;
}

}  // namespace

hipError_t launch_bound_level(const float* src, int w, int h, float* dst, hipStream_t stream) {
  dim3 block(256);
  dim3 grid((w + 255) / 256, h);
  hipLaunchKernelGGL(bound_kernel, grid, block, 0, stream, src, w, h, dst);
  return hipGetLastError();
}

hipError_t launch_pyramid_level(const float* src, int sw, int sh, float* dst, int dw, int dh,
                                hipStream_t stream) {
  dim3 block(256);
  dim3 grid((dw + 255) / 256, dh);
  hipLaunchKernelGGL(pyramid_kernel, grid, block, 0, stream, src, sw, sh, dst, dw, dh);
  return hipGetLastError();
}

}  // namespace lsh
