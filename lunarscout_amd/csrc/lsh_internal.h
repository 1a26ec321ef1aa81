// lsh_internal.h -- shared types of the HIP library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lunarscout_hip.h"

namespace lsh {

constexpr int kMaxDems = 4;        // DEM passes fused in one march launch
constexpr int kMaxLevels = 16;     // pyramid levels (a 4^15-pixel DEM)
constexpr int kSegmentFields = 18; // SEGMENT_FIELDS (contract.py:33-38)

// Thread-local error text set by every failing entry point.
void set_error(const std::string& message);

}  // namespace lsh

// Status helpers of the C ABI entry points.
#define LSH_CHECK_HIP(expr, what)                                              \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      lsh::set_error(std::string(what) + ": " + hipGetErrorString(_e));        \
      return _e == hipErrorOutOfMemory ? LSH_ENOMEM : LSH_EDEVICE;             \
    }                                                                          \
  } while (0)

#define LSH_REQUIRE(cond, message) \
  do {                             \
    if (!(cond)) {                 \
      lsh::set_error(message);     \
      return LSH_EINVAL;           \
    }                              \
  } while (0)

namespace lsh {

struct Dem {
  bool live = false;
  int32_t width = 0, height = 0;
  double geo_transform[6] = {0};
  double projection[6] = {0};
  float* level0 = nullptr;  // device float32[h][w], followed in the same allocation by
  float* mips = nullptr;    // float32[mip_count] (levels 1..L-1) = level0 + w * h, then
  float* bounds = nullptr;  // the 2x2 bound of level 0 and the mips, same layout
  int64_t bound_offset = 0;  // bounds - level0 (elements)
  int64_t mip_count = 0;
  int32_t level_count = 0;
  int32_t levels[kMaxLevels][4] = {{0}};  // (level, offset, width, height)
  int32_t all_valid = 0;  // every elevation finite and above -20000 m
};

// Per-pass description of a DEM as the march kernel reads it.
struct MarchDem {
  const float* level0;      // level 0, followed by the mips and the bounds (one allocation)
  const float* mips;
  uint32_t buffer_bytes;    // bytes of level 0 + mips + bounds
  int32_t bound_offset;     // element offset of the bounds from level0
  int32_t width, height, level_count;
  float map_resolution;     // (column + row resolution) / 2 (cuda_backend.py:1124-1127)
  float column_resolution;  // hypot(t1, t4) (cuda_backend.py:1032-1035)
  float radius;             // projection radius (cuda_backend.py:1132)
  int32_t all_valid;
  double resolution;        // map_resolution widened
  int32_t resolution_exponent;  // its binary exponent (frexp)
  double resolution_mantissa;   // and its mantissa in [0.5, 1)
  uint32_t box_offset;          // element offset of the level-2 bounds from level0
  int32_t box_width;            // level 2's width in cells; 0 when the DEM has no level 2
  int32_t level_offset[kMaxLevels];  // element offsets from level0 (level 0: 0)
  int32_t level_width[kMaxLevels];
  int32_t level_height[kMaxLevels];
};

struct MarchArgs {
  const float* segments;  // [N][C][D][18]
  float* output;          // [tile_h * tile_w][N]
  int32_t* winners;       // nullable [tile_h * tile_w][N] winning-sample index
  unsigned long long* counters;  // nullable [3]
  int32_t N, C, D;
  int32_t tile_column, tile_row, tile_width, tile_height, subpatch;
  float observer_elevation;
  int32_t flags;
  MarchDem dem[kMaxDems];
};

struct Session {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mutex;
  std::vector<Dem> dems;
  // session-owned scratch for the one-call tile path
  float* segment_scratch = nullptr;
  size_t segment_scratch_bytes = 0;
  lsh_tile* tile_scratch = nullptr;  // device copy of a call's tile list
  int32_t tile_scratch_count = 0;
  void* table_scratch = nullptr;     // device copy of a call's small tables
  size_t table_scratch_bytes = 0;
  int cu_count = 0;
};

// Sun-disk slice table (lightmap.py:15-29) for a half angle in degrees.
void sun_disk(double half_angle_deg, float* half_circle, float* max_photons);

// Copies a host tile list into session scratch (grow-only, stream ordered).
int upload_tiles(Session* s, const lsh_tile* tiles, int32_t count, const lsh_tile** device_tiles);
// Copies `bytes` of host tables into session scratch (grow-only, stream ordered).
int upload_tables(Session* s, const void* host, size_t bytes, void** device);

// Kernel launchers (defined in the .hip files).
hipError_t launch_bound_level(const float* src, int w, int h, float* dst, hipStream_t stream);
hipError_t launch_pyramid_level(const float* src, int sw, int sh, float* dst, int dw, int dh,
                                hipStream_t stream);

struct SegmentDem {
  const float* elevation;  // device level 0
  int32_t width, height;
  double t[6];
  double p[6];
  double map_resolution;
  double ray_limit;
};

struct SegmentArgs {
  float* out;  // [N][C][D][18]
  int32_t N, C, D, count;  // count = tile_width/subpatch + 2
  int32_t tile_column, tile_row, subpatch;
  double maximum_distance;
  double observer_elevation;
  SegmentDem dem[kMaxDems];
};

hipError_t launch_segments(const SegmentArgs& args, hipStream_t stream);

// SegmentDem of a registered DEM: map resolution and ray limit
// min(maximum distance, 1.2 * min(width, height) * resolution) (geometry.py:452-462).
SegmentDem segment_dem(const Dem& dem, double maximum_distance_m);
Dem* find_dem(Session* s, int32_t id);
int select_device(Session* s);

// Batches: a site is one DEM list (primary first) with its observer height; a
// batch tile (lsh_batch_tile) names its site and origin.
struct SegmentSite {
  int32_t D;
  double observer_elevation;
  SegmentDem dem[kMaxDems];
};

struct SegmentBatchArgs {
  const SegmentSite* sites;    // device
  const lsh_batch_tile* tiles; // device [tile_count]
  float* segments;             // device, tile slot i at segments + i * slot_floats
  int64_t slot_floats;
  int32_t tile_count, N, C, count, subpatch;
  double maximum_distance;
};

hipError_t launch_segments_batch(const SegmentBatchArgs& args, hipStream_t stream);

struct MarchSite {
  int32_t D;
  float observer_elevation;
  MarchDem dem[kMaxDems];
};

struct BatchMarchArgs {
  const MarchSite* sites;       // device
  const lsh_batch_tile* tiles;  // device [tile_count]
  const float* segments;        // device, tile slot i at segments + i * slot_floats
  int64_t slot_floats;
  float* output;                // device [tile][16384][N]
  unsigned int* queues;         // device [8] work counters, zero at launch
  int32_t tile_count, N, C, subpatch, flags;
};

// Kernel shapes of the batched march (pixel block kBX x kBY per wavefront,
// kAZ azimuths per workgroup); index 0 is the plain kernel's shape.
struct BatchShape {
  int32_t block_x, block_y, azimuths;
};
int batch_shape_count();
BatchShape batch_shape(int config);
// Launches the persistent batched march with `workgroups` workgroups (a
// multiple of 8); *resident receives the occupancy-derived default when
// workgroups <= 0 is passed.
hipError_t launch_march_batch(const BatchMarchArgs& args, int config, int workgroups,
                              hipStream_t stream);
int batch_default_workgroups(int config, int cu_count);
hipError_t launch_reset_queues(unsigned int* queues, hipStream_t stream);

struct LightArgs {
  const float* elevation;  // device level 0 of the DEM the tiles belong to
  int32_t dem_width, dem_height;
  double geo_transform[6];
  double projection[6];
  const lsh_tile* tiles;   // device [tile_count]
  int32_t tile_count;
  const float* horizons;   // device [tile][pixel][N] or [tile][N][pixel], degrees
  int32_t N;
  int32_t azimuth_major;
  const double* sun;       // device [sun_count][3]
  int32_t sun_count;
  const double* earth;     // device [earth_count][3]
  int32_t earth_count;
  const double* psr_vectors;  // device [psr_count][3]
  int32_t psr_count;
  float half_circle[16];   // _HALF_CIRCLE (lightmap.py:19-25)
  float max_photons;       // _MAX_PHOTONS (lightmap.py:26-29)
  float sun_half_angle;    // SUN_HALF_ANGLE_DEG (lightmap.py:15)
  double psr_threshold;    // -diameter / 2 (psr.py:115)
  uint8_t* sun_bytes;      // [tile][sun_count][16384]
  float* sun_fraction;     // [tile][sun_count][16384]
  float* sun_mean;         // [tile][16384]
  float* earth_margin;     // [tile][earth_count][16384]
  float* earth_visible;    // [tile][16384]
  uint8_t* psr;            // [tile][16384]
  float* psr_margin;       // [tile][16384]
};

hipError_t launch_lighting(const LightArgs& args, hipStream_t stream);
void run_lighting_host(const LightArgs& args, int threads);
hipError_t launch_march(const MarchArgs& args, hipStream_t stream);
hipError_t launch_sun_loss(const LightArgs& light, double smoothing, double* loss, float* grad,
                           hipStream_t stream);
hipError_t launch_adjoint(const MarchArgs& march, const float* grad_horizon,
                          unsigned long long* const* grad_fixed, double fixed_scale,
                          int32_t* sample_cells, hipStream_t stream);
hipError_t launch_probe_copy(const float* src, float* dst, int64_t bytes, hipStream_t stream);
hipError_t launch_probe_gather(const float* window, int32_t width, int32_t height, int32_t taps,
                               int32_t threads, float* out, hipStream_t stream);

// Fills a MarchDem from a registered DEM (float32 parameters the kernel reads).
MarchDem march_dem(const Dem& dem);

}  // namespace lsh
