// lsh_march.hip -- the horizon ray march of one tile, every DEM pass fused.
//
// Restates subpatch_hierarchy_kernel (cuda_backend.py:1007-1134) with its device
// helpers and the production march hierarchy_maximum (cuda_backend.py:656-874).
// The step schedule is the reference's: start level from the beam footprint,
// conservative 2x2 bound tests against the running horizon, nodata skips, the
// adaptive level-0 advance (pixel / margin / angular terms, near-field x0.25,
// pass-0 far floor) and no cell-exit cap unless LSH_HORIZON_CAP asks for the
// diagnostic schedule.
//
// MI355X mapping:
//   * a wavefront is an 8x8 pixel block at one azimuth: its 64 rays are parallel
//     and at most 8 px apart, so they visit the same pyramid cells and have
//     similar trip counts;
//   * a 256-thread workgroup is that pixel block at 4 consecutive azimuths;
//   * workgroup ids are remapped so each of the 8 XCDs owns a contiguous range
//     of azimuths (one L2 per XCD sees one angular sector of the DEM);
//   * every DEM pass runs in the same thread, the running maximum carried in a
//     register (the reference launches one kernel per pass on one buffer);
//   * the march is one flat loop: each iteration is one level step of the
//     hierarchy for every lane, so lanes at different levels stay converged.
//
// Arithmetic: float64 for the ray geometry and, by default, for heights and
// slopes too (DEM storage stays float32); LSH_HORIZON_F32_VALUES selects
// float32 heights and slopes (see lsh_march_core.h for the trade-off).

#include "lsh_internal.h"
#include "lsh_march_core.h"

namespace lsh {
namespace {

constexpr int kBlockPixels = 8;       // 8x8 pixel block per wavefront
constexpr int kAzimuthsPerGroup = 4;  // wavefronts per workgroup

// The traversal counters are always accumulated: besides feeding the
// roofline, their per-branch updates keep the compiler from splitting the flat
// march loop into nested divergent loops (that restructured form ran 9x slower
// on gfx950 with identical results). They are stored only when requested.
// The march is latency-bound on its dependent L2 gathers: capping registers
// for 4 wavefronts per SIMD (a few spills) measured 1.37x faster than the
// uncapped 2 wavefronts.
#ifndef LSH_MARCH_WAVES_PER_EU
#define LSH_MARCH_WAVES_PER_EU 4
#endif

template <class V, bool kCap>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LSH_MARCH_WAVES_PER_EU)))
march_kernel(MarchArgs args) {
  __shared__ int32_t s_off[kMaxDems][kMaxLevels];
  __shared__ int32_t s_w[kMaxDems][kMaxLevels];
  __shared__ int32_t s_h[kMaxDems][kMaxLevels];
  for (int i = threadIdx.x; i < kMaxDems * kMaxLevels; i += blockDim.x) {
    int dd = i / kMaxLevels, l = i % kMaxLevels;
    s_off[dd][l] = args.dem[dd].level_offset[l];
    s_w[dd][l] = args.dem[dd].level_width[l];
    s_h[dd][l] = args.dem[dd].level_height[l];
  }
  __syncthreads();

  // XCD-aware workgroup order: hardware deals workgroup ids round-robin to the
  // 8 XCDs; give XCD k a contiguous range of (azimuth group, pixel block) items.
  const int nbx = (args.tile_width + kBlockPixels - 1) / kBlockPixels;
  const int nby = (args.tile_height + kBlockPixels - 1) / kBlockPixels;
  const int blocks_per_group = nbx * nby;
  const int total = gridDim.x;
  const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const int base = total / 8, extra = total % 8;
  const int logical = xcd * base + min(xcd, extra) + slot;
  const int group = logical / blocks_per_group;
  const int pblock = logical % blocks_per_group;

  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int az = group * kAzimuthsPerGroup + wave;
  const int column = (pblock % nbx) * kBlockPixels + lane % kBlockPixels;
  const int row = (pblock / nbx) * kBlockPixels + lane / kBlockPixels;
  if (az >= args.N || column >= args.tile_width || row >= args.tile_height) return;
  const int pixel = row * args.tile_width + column;
  const int64_t out_index = static_cast<int64_t>(pixel) * args.N + az;
  const MarchDem& primary = args.dem[0];
  if (args.tile_column + column >= primary.width || args.tile_row + row >= primary.height) {
    args.output[out_index] = -INFINITY;  // padded pixels keep the sentinel
    if (args.winners) args.winners[out_index] = -1;
    return;
  }
  const PixelSelection sel = lsh_select(column, row, args.tile_column, args.tile_row,
                                        args.tile_width, args.subpatch, primary.width,
                                        primary.height);
  const V oz = lsh_observer_terrain<V>(primary.level0, primary.width, primary.height,
                                       args.tile_column + column, args.tile_row + row) +
               static_cast<V>(args.observer_elevation);
  V current = V(-1e30);
  int32_t best = -1;
  MarchCounters cnt{0, 0, 0};
  const int64_t seg_stride = static_cast<int64_t>(args.D) * kSegmentFields;
  const float* seg_az = args.segments + static_cast<int64_t>(az) * args.C * seg_stride;
  for (int pass = 0; pass < args.D; ++pass) {
    const MarchDem& d = args.dem[pass];
    const float* p = seg_az + pass * kSegmentFields;
    const PassSegment<double> g = lsh_interpolate<double>(p + sel.i00 * seg_stride, p + sel.i10 * seg_stride,
                                          p + sel.i01 * seg_stride, p + sel.i11 * seg_stride,
                                          sel, column, row,
                                          primary.column_resolution / d.column_resolution);
    MarchView view{d.level0, d.mips, s_off[pass], s_w[pass], s_h[pass], d.width, d.height,
                   d.level_count, d.map_resolution, d.radius, d.all_valid};
    lsh_march_pass<double, V, kCap, true>(g, view, oz, pass, current, best, cnt);
  }
  float result = current > V(-1e29) ? static_cast<float>(current) : -INFINITY;
  if (args.flags & LSH_HORIZON_DEGREES) {
    // HorizonBuffers.degrees (contract.py:362-371)
    result = isinf(result) ? result : atanf(result) * (180.0f / 3.14159265358979323846f);
  }
  args.output[out_index] = result;
  if (args.winners) args.winners[out_index] = best;
  if (args.counters) {
    // counters[3]: 64 x the wavefront's longest lane (iterations the SIMD issues)
    unsigned long long a = cnt.samples, b = cnt.tests, c = cnt.steps;
    unsigned m = cnt.steps;
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_down(a, o);
      b += __shfl_down(b, o);
      c += __shfl_down(c, o);
      unsigned other = __shfl_xor(m, o);
      m = other > m ? other : m;
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&args.counters[0], a);
      atomicAdd(&args.counters[1], b);
      atomicAdd(&args.counters[2], c);
      atomicAdd(&args.counters[3], 64ull * m);
    }
  }
}

}  // namespace

MarchDem march_dem(const Dem& dem) {
  MarchDem m{};
  m.level0 = dem.level0;
  m.mips = dem.mips;
  m.width = dem.width;
  m.height = dem.height;
  m.level_count = dem.level_count;
  // float32 map parameters as the reference kernel reads them (pyramid.py:78-107)
  float t1 = static_cast<float>(dem.geo_transform[1]), t4 = static_cast<float>(dem.geo_transform[4]);
  float t2 = static_cast<float>(dem.geo_transform[2]), t5 = static_cast<float>(dem.geo_transform[5]);
  float column_res = std::sqrt(t1 * t1 + t4 * t4);
  float row_res = std::sqrt(t2 * t2 + t5 * t5);
  m.column_resolution = column_res;
  m.map_resolution = (column_res + row_res) * 0.5f;
  m.radius = static_cast<float>(dem.projection[0]);
  m.all_valid = dem.all_valid;
  for (int l = 0; l < kMaxLevels; ++l) {
    bool used = l < dem.level_count;
    m.level_offset[l] = used ? dem.levels[l][1] : 0;
    m.level_width[l] = used ? dem.levels[l][2] : 1;
    m.level_height[l] = used ? dem.levels[l][3] : 1;
  }
  return m;
}

template <class V>
void launch_march_values(MarchArgs args, dim3 grid, dim3 block, hipStream_t stream) {
  if (!(args.flags & LSH_HORIZON_COUNTERS)) args.counters = nullptr;
  if (args.flags & LSH_HORIZON_CAP)
    hipLaunchKernelGGL((march_kernel<V, true>), grid, block, 0, stream, args);
  else
    hipLaunchKernelGGL((march_kernel<V, false>), grid, block, 0, stream, args);
}

hipError_t launch_march(const MarchArgs& args, hipStream_t stream) {
  int nbx = (args.tile_width + kBlockPixels - 1) / kBlockPixels;
  int nby = (args.tile_height + kBlockPixels - 1) / kBlockPixels;
  int groups = (args.N + kAzimuthsPerGroup - 1) / kAzimuthsPerGroup;
  dim3 grid(nbx * nby * groups), block(64 * kAzimuthsPerGroup);
  if (args.flags & LSH_HORIZON_F32_VALUES)
    launch_march_values<float>(args, grid, block, stream);
  else
    launch_march_values<double>(args, grid, block, stream);
  return hipGetLastError();
}

}  // namespace lsh
