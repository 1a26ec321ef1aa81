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

// Registers are sized for 4 wavefronts per SIMD (128 VGPRs). The level test
// forms the cell exit while its bound load is in flight, which keeps more
// values live; at 4 wavefronts the plain kernel spills 44 B per lane and the
// batched one 100 B. On the full-site tiles this measured 367 ms per tile
// against 379 ms at 5 wavefronts (96 VGPRs, 152 B of spills; the batched
// kernel 1.6x slower there); the uncapped build holds 2.
#ifndef LSH_MARCH_WAVES_PER_EU
#define LSH_MARCH_WAVES_PER_EU 4
#endif

// One ray fan of one pixel at one azimuth over every DEM pass, written to
// output[pixel * N + az]. `dems` are the tile's D DEMs (kernel arguments or
// the batch site table); s_off / s_w / s_h their level tables in LDS. kCount
// builds the traversal counters (a separate instantiation: counting costs
// registers and instructions in every iteration).
template <class V, bool kCap, bool kCount>
__device__ __attribute__((always_inline)) void march_pixel(
    const MarchDem* dems, const int32_t (*s_off)[kMaxLevels], const int32_t (*s_w)[kMaxLevels],
    const int32_t (*s_h)[kMaxLevels], int D, const float* segments, int N, int C,
    int tile_column, int tile_row, int tile_width, int tile_height, int subpatch,
    float observer_elevation, int flags, int column, int row, int az, float* output,
    int32_t* winners, unsigned long long* counters) {
  const int pixel = row * tile_width + column;
  // pixel-major (the tile file order) or azimuth-major: a wavefront (one
  // azimuth, 8 rows of 8 pixels) then stores 8 contiguous 32-byte runs
  const int64_t out_index = (flags & LSH_HORIZON_AZIMUTH_MAJOR)
                                ? static_cast<int64_t>(az) * (tile_width * tile_height) + pixel
                                : static_cast<int64_t>(pixel) * N + az;
  const MarchDem& primary = dems[0];
  if (tile_column + column >= primary.width || tile_row + row >= primary.height) {
    output[out_index] = -INFINITY;  // padded pixels keep the sentinel
    if (winners) winners[out_index] = -1;
    return;
  }
  const PixelSelection sel = lsh_select(column, row, tile_column, tile_row, tile_width, subpatch,
                                        primary.width, primary.height);
  const V oz = lsh_observer_terrain<V>(primary.level0, primary.width, primary.height,
                                       tile_column + column, tile_row + row) +
               static_cast<V>(observer_elevation);
  V current = V(-1e30);
  int32_t best = -1;
  MarchCounters cnt{0, 0, 0};
  const int seg_stride = D * kSegmentFields;
  const float* seg_az = segments + static_cast<int64_t>(az) * C * seg_stride;
  for (int pass = 0; pass < D; ++pass) {
    const MarchDem& d = dems[pass];
    const float* p = seg_az + pass * kSegmentFields;
    // the selection as opaque copies: what is derived from it is formed again
    // in each pass (a few instructions) instead of being held across the march,
    // where registers are short (held values spill)
    PixelSelection ps = sel;
    lsh_keep_value(ps.i00);
    lsh_keep_value(ps.i10);
    lsh_keep_value(ps.i01);
    lsh_keep_value(ps.i11);
    lsh_keep_value(ps.cl);
    lsh_keep_value(ps.cr);
    lsh_keep_value(ps.ct);
    lsh_keep_value(ps.cb);
    const PassSegment<double> g = lsh_interpolate<double>(
        p + ps.i00 * seg_stride, p + ps.i10 * seg_stride, p + ps.i01 * seg_stride,
        p + ps.i11 * seg_stride, ps, column, row,
        primary.column_resolution / d.column_resolution);
    MarchView view{d.level0, d.buffer_bytes, d.bound_offset, s_off[pass], s_w[pass], s_h[pass],
                   d.width, d.height, d.level_count, d.map_resolution, d.radius, d.all_valid,
                   d.resolution, d.resolution_exponent, d.resolution_mantissa,
                   d.box_offset, d.box_width};
    lsh_march_pass<double, V, kCap, kCount>(g, view, oz, pass, current, best, cnt);
  }
  float result = current > V(-1e29) ? static_cast<float>(current) : -INFINITY;
  if (flags & LSH_HORIZON_DEGREES) {
    // HorizonBuffers.degrees (contract.py:362-371)
    result = isinf(result) ? result : atanf(result) * (180.0f / 3.14159265358979323846f);
  }
  output[out_index] = result;
  if (winners) winners[out_index] = best;
  if (kCount && counters) {
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
      atomicAdd(&counters[0], a);
      atomicAdd(&counters[1], b);
      atomicAdd(&counters[2], c);
      atomicAdd(&counters[3], 64ull * m);
    }
  }
}

template <class V, bool kCap, bool kCount>
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
  march_pixel<V, kCap, kCount>(args.dem, s_off, s_w, s_h, args.D, args.segments, args.N,
                               args.C, args.tile_column, args.tile_row, args.tile_width,
                               args.tile_height, args.subpatch, args.observer_elevation,
                               args.flags, column, row, az, args.output, args.winners,
                               args.counters);
}

// The persistent batched march. Work items are (azimuth group, tile, pixel
// block) in that order; XCD k owns queue k, the k-th eighth of the items (one
// angular sector of every tile in the batch, so its L2 holds one sector of each
// DEM), and steals from the other queues once its own is drained. Thread 0
// takes an item with one atomic add and hands it to the workgroup through LDS
// (double-buffered by iteration parity, so no second barrier is needed); a
// workgroup leaves when every queue is exhausted, which each reaches since the
// counters only grow.
template <int kBX, int kBY, int kAZ>
__global__ void __launch_bounds__(64 * kAZ) __attribute__((amdgpu_waves_per_eu(LSH_MARCH_WAVES_PER_EU)))
march_batch_kernel(BatchMarchArgs a) {
  static_assert(kBX * kBY == 64, "a wavefront holds 64 pixels");
  __shared__ int32_t s_off[kMaxDems][kMaxLevels];
  __shared__ int32_t s_w[kMaxDems][kMaxLevels];
  __shared__ int32_t s_h[kMaxDems][kMaxLevels];
  __shared__ int32_t s_item[2];
  constexpr int kNbx = 128 / kBX, kNby = 128 / kBY, kBlocks = kNbx * kNby;
  const int groups = (a.N + kAZ - 1) / kAZ;
  const int64_t per_group = static_cast<int64_t>(a.tile_count) * kBlocks;
  const int64_t total = per_group * groups;
  const int xcd = blockIdx.x % 8;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int loaded_site = -1;
  for (int iteration = 0;; ++iteration) {
    const int parity = iteration & 1;
    if (threadIdx.x == 0) {
      int64_t item = -1;
      for (int k = 0; k < 8 && item < 0; ++k) {
        const int q = (xcd + k) & 7;
        const int64_t begin = total * q / 8, end = total * (q + 1) / 8;
        if (begin >= end) continue;
        const unsigned got = atomicAdd(&a.queues[q], 1u);
        if (begin + static_cast<int64_t>(got) < end) item = begin + got;
      }
      s_item[parity] = static_cast<int32_t>(item);
    }
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item[parity]);
    if (item < 0) break;
    const int group = static_cast<int>(item / per_group);
    const int rest = static_cast<int>(item % per_group);
    const int tile = rest / kBlocks, pblock = rest % kBlocks;
    const lsh_batch_tile t = a.tiles[tile];
    const int site = __builtin_amdgcn_readfirstlane(t.site);
    const MarchSite& ms = a.sites[site];
    if (site != loaded_site) {  // every wave left the previous item at the barrier above
      for (int i = threadIdx.x; i < kMaxDems * kMaxLevels; i += blockDim.x) {
        const int dd = i / kMaxLevels, l = i % kMaxLevels;
        s_off[dd][l] = ms.dem[dd].level_offset[l];
        s_w[dd][l] = ms.dem[dd].level_width[l];
        s_h[dd][l] = ms.dem[dd].level_height[l];
      }
      __syncthreads();
      loaded_site = site;
    }
    const int az = group * kAZ + wave;
    const int column = (pblock % kNbx) * kBX + lane % kBX;
    const int row = (pblock / kNbx) * kBY + lane / kBX;
    if (az < a.N) {
      march_pixel<double, false, false>(ms.dem, s_off, s_w, s_h, ms.D,
                                 a.segments + tile * a.slot_floats, a.N, a.C, t.tile_column,
                                 t.tile_row, 128, 128, a.subpatch, ms.observer_elevation,
                                 a.flags, column, row, az,
                                 a.output + static_cast<int64_t>(tile) * 128 * 128 * a.N,
                                 nullptr, nullptr);
    }
  }
}

__global__ void reset_queues_kernel(unsigned int* queues) {
  if (threadIdx.x < 8) queues[threadIdx.x] = 0u;
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
  m.resolution = static_cast<double>(m.map_resolution);
  m.resolution_mantissa = std::frexp(m.resolution, &m.resolution_exponent);
  m.radius = static_cast<float>(dem.projection[0]);
  m.all_valid = dem.all_valid;
  const int64_t level0_floats = static_cast<int64_t>(dem.width) * dem.height;
  m.bound_offset = static_cast<int32_t>(dem.bound_offset);
  m.buffer_bytes = static_cast<uint32_t>(8 * dem.bound_offset);
  for (int l = 0; l < kMaxLevels; ++l) {
    bool used = l < dem.level_count;
    m.level_offset[l] = used && l > 0 ? static_cast<int32_t>(level0_floats + dem.levels[l][1]) : 0;
    m.level_width[l] = used ? dem.levels[l][2] : 1;
    m.level_height[l] = used ? dem.levels[l][3] : 1;
  }
  // the guarded skips read the level-2 bounds (lsh_march_core.h)
  m.box_offset = dem.level_count > 2
                     ? static_cast<uint32_t>(m.bound_offset + m.level_offset[2])
                     : 0u;
  m.box_width = dem.level_count > 2 ? m.level_width[2] : 0;
  return m;
}

template <class V>
void launch_march_values(MarchArgs args, dim3 grid, dim3 block, hipStream_t stream) {
  if (!(args.flags & LSH_HORIZON_COUNTERS)) args.counters = nullptr;
  const bool cap = args.flags & LSH_HORIZON_CAP, count = args.counters != nullptr;
  if (cap && count)
    hipLaunchKernelGGL((march_kernel<V, true, true>), grid, block, 0, stream, args);
  else if (cap)
    hipLaunchKernelGGL((march_kernel<V, true, false>), grid, block, 0, stream, args);
  else if (count)
    hipLaunchKernelGGL((march_kernel<V, false, true>), grid, block, 0, stream, args);
  else
    hipLaunchKernelGGL((march_kernel<V, false, false>), grid, block, 0, stream, args);
}

namespace {

constexpr BatchShape kShapes[] = {{8, 8, 4}, {8, 8, 2}, {16, 4, 4}, {4, 16, 4},
                                  {16, 4, 2}, {4, 16, 2}, {8, 8, 1}};

template <int kBX, int kBY, int kAZ>
const void* batch_kernel_symbol() {
  return reinterpret_cast<const void*>(&march_batch_kernel<kBX, kBY, kAZ>);
}

const void* batch_kernel(int config) {
  switch (config) {
    case 0: return batch_kernel_symbol<8, 8, 4>();
    case 1: return batch_kernel_symbol<8, 8, 2>();
    case 2: return batch_kernel_symbol<16, 4, 4>();
    case 3: return batch_kernel_symbol<4, 16, 4>();
    case 4: return batch_kernel_symbol<16, 4, 2>();
    case 5: return batch_kernel_symbol<4, 16, 2>();
    default: return batch_kernel_symbol<8, 8, 1>();
  }
}

}  // namespace

int batch_shape_count() { return static_cast<int>(sizeof(kShapes) / sizeof(kShapes[0])); }

BatchShape batch_shape(int config) { return kShapes[config]; }

int batch_default_workgroups(int config, int cu_count) {
  int per_cu = 0;
  const BatchShape shape = kShapes[config];
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel(config),
                                                   64 * shape.azimuths, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 16 / shape.azimuths;
  const int wgs = per_cu * (cu_count > 0 ? cu_count : 256);
  return wgs < 8 ? 8 : wgs - wgs % 8;
}

hipError_t launch_march_batch(const BatchMarchArgs& args, int config, int workgroups,
                              hipStream_t stream) {
  const BatchShape shape = kShapes[config];
  void* params[] = {const_cast<BatchMarchArgs*>(&args)};
  return hipLaunchKernel(batch_kernel(config), dim3(workgroups), dim3(64 * shape.azimuths),
                         params, 0, stream);
}

hipError_t launch_reset_queues(unsigned int* queues, hipStream_t stream) {
  hipLaunchKernelGGL(reset_queues_kernel, dim3(1), dim3(64), 0, stream, queues);
  return hipGetLastError();
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
