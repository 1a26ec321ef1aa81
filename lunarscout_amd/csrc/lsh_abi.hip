// lsh_abi.hip -- the C ABI: sessions, device memory, DEM registration and the
// horizon entry points. Declarations and reference citations are in
// include/lunarscout_hip.h.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lsh_internal.h"

namespace lsh {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& message) { g_error = message; }

}  // namespace lsh

using lsh::Dem;
using lsh::Session;
using lsh::set_error;

namespace {

int pyramid_layout(int w, int h, int32_t (*levels)[4], int cap, int64_t* mip_count) {
  int count = 0;
  int64_t offset = 0;
  int cw = w, ch = h;
  while (true) {
    if (count >= cap) return -1;
    levels[count][0] = count;
    levels[count][1] = count == 0 ? 0 : static_cast<int32_t>(offset);
    levels[count][2] = cw;
    levels[count][3] = ch;
    if (count > 0) offset += static_cast<int64_t>(cw) * ch;
    ++count;
    if (cw == 1 && ch == 1) break;
    cw = (cw + 3) / 4;
    ch = (ch + 3) / 4;
  }
  *mip_count = offset;
  return count;
}

}  // namespace

namespace lsh {

Dem* find_dem(Session* s, int32_t id) {
  if (!s || id < 0 || id >= static_cast<int32_t>(s->dems.size()) || !s->dems[id].live) return nullptr;
  return &s->dems[id];
}

int select_device(Session* s) {
  LSH_CHECK_HIP(hipSetDevice(s->device), "hipSetDevice");
  return LSH_OK;
}

}  // namespace lsh

using lsh::find_dem;
using lsh::select_device;

namespace lsh {

void sun_disk(double half_angle_deg, float* half_circle, float* max_photons) {
  float angle = static_cast<float>(half_angle_deg);
  float sum = 0.0f;
  for (int i = 0; i < 16; ++i) {
    half_circle[i] = static_cast<float>(std::sqrt(64.0 - (7.5 - i) * (7.5 - i)) / 8.0 *
                                        static_cast<double>(angle));
    sum = sum + half_circle[i];
  }
  *max_photons = 2.0f * sum;
}

int upload_tiles(Session* s, const lsh_tile* tiles, int32_t count, const lsh_tile** device_tiles) {
  if (s->tile_scratch_count < count) {
    LSH_CHECK_HIP(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    if (s->tile_scratch) hipFree(s->tile_scratch);
    s->tile_scratch = nullptr;
    s->tile_scratch_count = 0;
    LSH_CHECK_HIP(hipMalloc(&s->tile_scratch, sizeof(lsh_tile) * count), "hipMalloc tile list");
    s->tile_scratch_count = count;
  }
  LSH_CHECK_HIP(hipMemcpyAsync(s->tile_scratch, tiles, sizeof(lsh_tile) * count,
                               hipMemcpyHostToDevice, s->stream),
                "copy tile list");
  *device_tiles = s->tile_scratch;
  return LSH_OK;
}

int upload_tables(Session* s, const void* host, size_t bytes, void** device) {
  if (s->table_scratch_bytes < bytes) {
    LSH_CHECK_HIP(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    if (s->table_scratch) hipFree(s->table_scratch);
    s->table_scratch = nullptr;
    s->table_scratch_bytes = 0;
    LSH_CHECK_HIP(hipMalloc(&s->table_scratch, bytes), "hipMalloc tables");
    s->table_scratch_bytes = bytes;
  }
  LSH_CHECK_HIP(hipMemcpyAsync(s->table_scratch, host, bytes, hipMemcpyHostToDevice, s->stream),
                "copy tables");
  *device = s->table_scratch;
  return LSH_OK;
}

SegmentDem segment_dem(const Dem& d, double maximum_distance_m) {
  SegmentDem g{};
  g.elevation = d.level0;
  g.width = d.width;
  g.height = d.height;
  std::memcpy(g.t, d.geo_transform, sizeof(g.t));
  std::memcpy(g.p, d.projection, sizeof(g.p));
  // DemGrid.map_resolution_m and the ray limit (geometry.py:78-83, 452-462)
  g.map_resolution = (std::hypot(g.t[1], g.t[4]) + std::hypot(g.t[2], g.t[5])) * 0.5;
  double size_m = std::fmin(d.width * g.map_resolution, d.height * g.map_resolution);
  g.ray_limit = std::fmin(maximum_distance_m, size_m * 1.2);
  return g;
}

}  // namespace lsh

extern "C" {

int lsh_last_error(char* buffer, size_t capacity) {
  const std::string& e = lsh::g_error;
  if (buffer && capacity) {
    size_t n = e.size() < capacity - 1 ? e.size() : capacity - 1;
    std::memcpy(buffer, e.data(), n);
    buffer[n] = 0;
  }
  return static_cast<int>(e.size());
}

const char* lsh_version(void) { return "0.1.0"; }

int lsh_device_count(int32_t* count) {
  LSH_REQUIRE(count, "count must not be null");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = e == hipSuccess ? n : 0;
  return LSH_OK;
}

int lsh_device_info(int32_t device, char* name, size_t capacity, int32_t* compute_units) {
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    set_error(std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    return LSH_ENODEV;
  }
  if (name && capacity) {
    std::snprintf(name, capacity, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (compute_units) *compute_units = prop.multiProcessorCount;
  return LSH_OK;
}

int lsh_session_create(int32_t device, lsh_session** session) {
  LSH_REQUIRE(session, "session must not be null");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    set_error("no HIP device is visible");
    return LSH_ENODEV;
  }
  if (device < 0 || device >= n) {
    set_error("HIP device index out of range");
    return LSH_ENODEV;
  }
  Session* s = new Session();
  s->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error(std::string("cannot initialise HIP device: ") + hipGetErrorString(e));
    delete s;
    return LSH_ENODEV;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) s->cu_count = prop.multiProcessorCount;
  *session = reinterpret_cast<lsh_session*>(s);
  return LSH_OK;
}

int lsh_session_destroy(lsh_session* session) {
  Session* s = reinterpret_cast<Session*>(session);
  if (!s) return LSH_OK;
  hipSetDevice(s->device);
  hipStreamSynchronize(s->stream);
  for (Dem& d : s->dems) {
    if (d.live) hipFree(d.level0);  // the mips and bounds share the allocation
  }
  if (s->segment_scratch) hipFree(s->segment_scratch);
  if (s->tile_scratch) hipFree(s->tile_scratch);
  if (s->table_scratch) hipFree(s->table_scratch);
  hipStreamDestroy(s->stream);
  delete s;
  return LSH_OK;
}

int lsh_alloc(lsh_session* session, size_t bytes, void** device_pointer) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && device_pointer, "invalid arguments");
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(hipMalloc(device_pointer, bytes ? bytes : 1), "hipMalloc");
  return LSH_OK;
}

int lsh_free(lsh_session* session, void* device_pointer) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s, "invalid session");
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  LSH_CHECK_HIP(hipFree(device_pointer), "hipFree");
  return LSH_OK;
}

int lsh_copy_to_device(lsh_session* session, void* dst, const void* src, size_t bytes) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && dst && src, "invalid arguments");
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream), "copy to device");
  LSH_CHECK_HIP(hipStreamSynchronize(s->stream), "copy to device");
  return LSH_OK;
}

int lsh_copy_to_host(lsh_session* session, void* dst, const void* src, size_t bytes) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && dst && src, "invalid arguments");
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream), "copy to host");
  LSH_CHECK_HIP(hipStreamSynchronize(s->stream), "copy to host");
  return LSH_OK;
}

int lsh_memset_device(lsh_session* session, void* dst, int value, size_t bytes) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && dst, "invalid arguments");
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(hipMemsetAsync(dst, value, bytes, s->stream), "hipMemsetAsync");
  return LSH_OK;
}

int lsh_synchronize(lsh_session* session) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s, "invalid session");
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(hipStreamSynchronize(s->stream), "kernel execution");
  return LSH_OK;
}

int lsh_event_create(lsh_session* session, void** event) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && event, "invalid arguments");
  if (int r = select_device(s)) return r;
  hipEvent_t e;
  LSH_CHECK_HIP(hipEventCreate(&e), "hipEventCreate");
  *event = e;
  return LSH_OK;
}

int lsh_event_record(lsh_session* session, void* event) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && event, "invalid arguments");
  LSH_CHECK_HIP(hipEventRecord(static_cast<hipEvent_t>(event), s->stream), "hipEventRecord");
  return LSH_OK;
}

int lsh_event_elapsed_ms(void* start, void* stop, float* milliseconds) {
  LSH_REQUIRE(start && stop && milliseconds, "invalid arguments");
  LSH_CHECK_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)), "hipEventSynchronize");
  LSH_CHECK_HIP(hipEventElapsedTime(milliseconds, static_cast<hipEvent_t>(start),
                                    static_cast<hipEvent_t>(stop)),
                "hipEventElapsedTime");
  return LSH_OK;
}

int lsh_event_destroy(void* event) {
  if (event) hipEventDestroy(static_cast<hipEvent_t>(event));
  return LSH_OK;
}

int lsh_pyramid_layout(int32_t width, int32_t height, int32_t* levels, int32_t capacity,
                       int32_t* level_count, int64_t* mip_count) {
  LSH_REQUIRE(width > 0 && height > 0, "DEM dimensions must be positive");
  LSH_REQUIRE(level_count && mip_count, "invalid arguments");
  int32_t table[lsh::kMaxLevels][4];
  int64_t mips = 0;
  int n = pyramid_layout(width, height, table, lsh::kMaxLevels, &mips);
  LSH_REQUIRE(n > 0, "DEM too large for the pyramid table");
  if (levels) {
    LSH_REQUIRE(capacity >= n, "level table capacity too small");
    std::memcpy(levels, table, sizeof(int32_t) * 4 * n);
  }
  *level_count = n;
  *mip_count = mips;
  return LSH_OK;
}

int lsh_dem_create(lsh_session* session, const float* elevation, int32_t width, int32_t height,
                   const double* geo_transform, const double* projection,
                   const float* mips_host, int32_t* dem_id) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && elevation && geo_transform && projection && dem_id, "invalid arguments");
  LSH_REQUIRE(width > 0 && height > 0, "DEM dimensions must be positive");
  double det = geo_transform[1] * geo_transform[5] - geo_transform[2] * geo_transform[4];
  LSH_REQUIRE(std::fabs(det) >= 1e-12, "DEM geotransform is singular");
  for (int i = 0; i < 6; ++i)
    LSH_REQUIRE(std::isfinite(geo_transform[i]) && std::isfinite(projection[i]),
                "DEM transform and projection must be finite");
  // row offsets are formed with 24-bit multiplies in the march
  LSH_REQUIRE(width < (1 << 24) && height < (1 << 24), "DEM dimensions must be below 2^24");
  Dem d;
  int n = pyramid_layout(width, height, d.levels, lsh::kMaxLevels, &d.mip_count);
  LSH_REQUIRE(n > 0, "DEM too large for the pyramid table");
  d.level_count = n;
  d.width = width;
  d.height = height;
  std::memcpy(d.geo_transform, geo_transform, sizeof(d.geo_transform));
  std::memcpy(d.projection, projection, sizeof(d.projection));
  d.all_valid = 1;
  for (int64_t i = 0, n = static_cast<int64_t>(width) * height; i < n; ++i) {
    float v = elevation[i];
    if (!(std::isfinite(v) && v > -20000.0f)) {
      d.all_valid = 0;
      break;
    }
  }
  if (int r = select_device(s)) return r;
  std::lock_guard<std::mutex> lock(s->mutex);
  // level 0, its mips and the bound copy of both in one allocation, so the march
  // addresses every level with 32-bit offsets from one base (buffer loads)
  size_t level0_bytes = sizeof(float) * static_cast<size_t>(width) * height;
  size_t mip_bytes = sizeof(float) * static_cast<size_t>(d.mip_count > 0 ? d.mip_count : 1);
  LSH_REQUIRE(2 * (level0_bytes + mip_bytes) < (static_cast<size_t>(1) << 32),
              "DEM, pyramid and bounds must fit in 4 GiB");
  LSH_CHECK_HIP(hipMalloc(&d.level0, 2 * (level0_bytes + mip_bytes)), "hipMalloc DEM");
  d.mips = d.level0 + static_cast<size_t>(width) * height;
  d.bound_offset = static_cast<int64_t>((level0_bytes + mip_bytes) / sizeof(float));
  d.bounds = d.level0 + d.bound_offset;
  hipError_t e = hipSuccess;
  e = hipMemcpyAsync(d.level0, elevation, level0_bytes, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess && mips_host && d.mip_count > 0)
    e = hipMemcpyAsync(d.mips, mips_host, sizeof(float) * d.mip_count, hipMemcpyHostToDevice,
                       s->stream);
  if (e == hipSuccess && !mips_host) {
    const float* src = d.level0;
    int sw = width, sh = height;
    for (int l = 1; l < n && e == hipSuccess; ++l) {
      float* dst = d.mips + d.levels[l][1];
      e = lsh::launch_pyramid_level(src, sw, sh, dst, d.levels[l][2], d.levels[l][3], s->stream);
      src = dst;
      sw = d.levels[l][2];
      sh = d.levels[l][3];
    }
  }
  for (int l = 0; l < n && e == hipSuccess; ++l) {
    const int64_t at = l == 0 ? 0 : static_cast<int64_t>(width) * height + d.levels[l][1];
    e = lsh::launch_bound_level(d.level0 + at, d.levels[l][2], d.levels[l][3], d.bounds + at,
                                s->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    hipFree(d.level0);
    set_error(std::string("DEM upload / pyramid / bounds build: ") + hipGetErrorString(e));
    return LSH_EDEVICE;
  }
  d.live = true;
  s->dems.push_back(d);
  *dem_id = static_cast<int32_t>(s->dems.size() - 1);
  return LSH_OK;
}

int lsh_dem_destroy(lsh_session* session, int32_t dem_id) {
  Session* s = reinterpret_cast<Session*>(session);
  Dem* d = find_dem(s, dem_id);
  LSH_REQUIRE(d, "unknown DEM id");
  if (int r = select_device(s)) return r;
  hipStreamSynchronize(s->stream);
  hipFree(d->level0);  // the mips and bounds share the allocation
  d->live = false;
  return LSH_OK;
}

int lsh_dem_pyramid(lsh_session* session, int32_t dem_id, float* mips_host, int64_t capacity) {
  Session* s = reinterpret_cast<Session*>(session);
  Dem* d = find_dem(s, dem_id);
  LSH_REQUIRE(d && mips_host, "invalid arguments");
  LSH_REQUIRE(capacity >= d->mip_count, "mip buffer too small");
  if (d->mip_count == 0) return LSH_OK;
  return lsh_copy_to_host(session, mips_host, d->mips, sizeof(float) * d->mip_count);
}

int lsh_dem_device_pointers(lsh_session* session, int32_t dem_id, const float** level0,
                            const float** mips) {
  Session* s = reinterpret_cast<Session*>(session);
  Dem* d = find_dem(s, dem_id);
  LSH_REQUIRE(d, "unknown DEM id");
  if (level0) *level0 = d->level0;
  if (mips) *mips = d->mips;
  return LSH_OK;
}

int lsh_segments_tile(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                      int32_t tile_column, int32_t tile_row, int32_t tile_width,
                      int32_t azimuth_count, double maximum_distance_m,
                      double observer_elevation_m, int32_t subpatch_size, float* segments) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && dem_ids && segments, "invalid arguments");
  LSH_REQUIRE(dem_count >= 1 && dem_count <= lsh::kMaxDems, "dem_count must be 1..4");
  LSH_REQUIRE(azimuth_count > 0, "azimuth_count must be positive");
  LSH_REQUIRE(subpatch_size > 0 && tile_width > 0, "tile width and subpatch size must be positive");
  LSH_REQUIRE(std::isfinite(maximum_distance_m) && std::isfinite(observer_elevation_m),
              "distances must be finite");
  lsh::SegmentArgs a{};
  a.out = segments;
  a.N = azimuth_count;
  a.count = tile_width / subpatch_size + 2;
  a.C = a.count * a.count;
  a.D = dem_count;
  a.tile_column = tile_column;
  a.tile_row = tile_row;
  a.subpatch = subpatch_size;
  a.maximum_distance = maximum_distance_m;
  a.observer_elevation = observer_elevation_m;
  for (int i = 0; i < dem_count; ++i) {
    Dem* d = find_dem(s, dem_ids[i]);
    LSH_REQUIRE(d, "unknown DEM id");
    a.dem[i] = lsh::segment_dem(*d, maximum_distance_m);
  }
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(lsh::launch_segments(a, s->stream), "segment kernel launch");
  return LSH_OK;
}

// Validates a tile march and fills MarchArgs (everything but output / winners).
static int march_args(Session* s, const int32_t* dem_ids, int32_t dem_count,
                      const float* segments, int32_t azimuth_count, int32_t tile_column,
                      int32_t tile_row, int32_t tile_width, int32_t tile_height,
                      int32_t subpatch_size, float observer_elevation_m, int32_t flags,
                      lsh::MarchArgs* out) {
  LSH_REQUIRE(s && dem_ids && segments, "invalid arguments");
  LSH_REQUIRE(dem_count >= 1 && dem_count <= lsh::kMaxDems, "dem_count must be 1..4");
  LSH_REQUIRE(azimuth_count > 0, "azimuth_count must be positive");
  LSH_REQUIRE(tile_width > 0 && tile_height > 0 && tile_width <= 128 && tile_height <= 128,
              "tile dimensions must be 1..128");
  LSH_REQUIRE(subpatch_size > 0 && tile_width / subpatch_size + 2 >= 2, "invalid subpatch size");
  LSH_REQUIRE(std::isfinite(observer_elevation_m), "observer elevation must be finite");
  lsh::MarchArgs& a = *out;
  a = lsh::MarchArgs{};
  a.segments = segments;
  a.N = azimuth_count;
  int count = tile_width / subpatch_size + 2;
  a.C = count * count;
  a.D = dem_count;
  a.tile_column = tile_column;
  a.tile_row = tile_row;
  a.tile_width = tile_width;
  a.tile_height = tile_height;
  a.subpatch = subpatch_size;
  a.observer_elevation = observer_elevation_m;
  a.flags = flags;
  for (int i = 0; i < dem_count; ++i) {
    Dem* d = find_dem(s, dem_ids[i]);
    LSH_REQUIRE(d, "unknown DEM id");
    a.dem[i] = lsh::march_dem(*d);
  }
  LSH_REQUIRE(tile_column >= 0 && tile_row >= 0 && tile_column < a.dem[0].width &&
                  tile_row < a.dem[0].height,
              "tile origin lies outside the primary DEM");
  return LSH_OK;
}

int lsh_horizon_tile(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                     const float* segments, int32_t azimuth_count, int32_t tile_column,
                     int32_t tile_row, int32_t tile_width, int32_t tile_height,
                     int32_t subpatch_size, float observer_elevation_m, int32_t flags,
                     float* output, int64_t* counters) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(output, "output must not be null");
  lsh::MarchArgs a;
  if (int r = march_args(s, dem_ids, dem_count, segments, azimuth_count, tile_column, tile_row,
                         tile_width, tile_height, subpatch_size, observer_elevation_m, flags, &a))
    return r;
  a.output = output;
  a.winners = nullptr;
  a.counters = reinterpret_cast<unsigned long long*>(counters);
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(lsh::launch_march(a, s->stream), "march kernel launch");
  return LSH_OK;
}

int lsh_horizon_forward(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                        const float* segments, int32_t azimuth_count, int32_t tile_column,
                        int32_t tile_row, float observer_elevation_m, int32_t flags,
                        float* output, int32_t* winners) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(output && winners, "output and winners must not be null");
  LSH_REQUIRE((flags & ~LSH_HORIZON_DEGREES) == 0, "forward flags: 0 or LSH_HORIZON_DEGREES");
  lsh::MarchArgs a;
  if (int r = march_args(s, dem_ids, dem_count, segments, azimuth_count, tile_column, tile_row,
                         128, 128, 8, observer_elevation_m, flags, &a))
    return r;
  a.output = output;
  a.winners = winners;
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(lsh::launch_march(a, s->stream), "march kernel launch");
  return LSH_OK;
}

int lsh_horizon_adjoint(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                        const float* segments, int32_t azimuth_count, int32_t tile_column,
                        int32_t tile_row, float observer_elevation_m, const int32_t* winners,
                        const float* grad_horizon, int64_t* const* grad_fixed,
                        double fixed_scale, int32_t* sample_cells) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(winners && grad_horizon && grad_fixed, "invalid arguments");
  LSH_REQUIRE(std::isfinite(fixed_scale) && fixed_scale > 0.0, "fixed_scale must be positive");
  lsh::MarchArgs a;
  if (int r = march_args(s, dem_ids, dem_count, segments, azimuth_count, tile_column, tile_row,
                         128, 128, 8, observer_elevation_m, 0, &a))
    return r;
  for (int i = 0; i < dem_count; ++i) LSH_REQUIRE(grad_fixed[i], "a gradient buffer is null");
  a.winners = const_cast<int32_t*>(winners);
  unsigned long long* grids[lsh::kMaxDems] = {nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < dem_count; ++i) grids[i] = reinterpret_cast<unsigned long long*>(grad_fixed[i]);
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(lsh::launch_adjoint(a, grad_horizon, grids, fixed_scale, sample_cells, s->stream),
                "adjoint kernel launch");
  return LSH_OK;
}

int lsh_horizon_tile_full(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                          int32_t tile_column, int32_t tile_row, int32_t azimuth_count,
                          double maximum_distance_m, float observer_elevation_m,
                          int32_t flags, float* output, int64_t* counters) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s, "invalid session");
  LSH_REQUIRE(dem_count >= 1 && dem_count <= lsh::kMaxDems, "dem_count must be 1..4");
  LSH_REQUIRE(azimuth_count > 0, "azimuth_count must be positive");
  size_t bytes = sizeof(float) * static_cast<size_t>(azimuth_count) * 324 * dem_count * 18;
  if (int r = select_device(s)) return r;
  if (s->segment_scratch_bytes < bytes) {
    LSH_CHECK_HIP(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    if (s->segment_scratch) hipFree(s->segment_scratch);
    s->segment_scratch = nullptr;
    s->segment_scratch_bytes = 0;
    LSH_CHECK_HIP(hipMalloc(&s->segment_scratch, bytes), "hipMalloc segment scratch");
    s->segment_scratch_bytes = bytes;
  }
  int r = lsh_segments_tile(session, dem_ids, dem_count, tile_column, tile_row, 128,
                            azimuth_count, maximum_distance_m, observer_elevation_m, 8,
                            s->segment_scratch);
  if (r) return r;
  return lsh_horizon_tile(session, dem_ids, dem_count, s->segment_scratch, azimuth_count,
                          tile_column, tile_row, 128, 128, 8, observer_elevation_m, flags,
                          output, counters);
}

int lsh_sun_disk(double sun_half_angle_deg, float* half_circle, float* max_photons) {
  LSH_REQUIRE(half_circle && max_photons, "invalid arguments");
  LSH_REQUIRE(std::isfinite(sun_half_angle_deg) && sun_half_angle_deg > 0.0,
              "sun half angle must be positive");
  lsh::sun_disk(sun_half_angle_deg, half_circle, max_photons);
  return LSH_OK;
}

// Validates a lighting call and fills everything of LightArgs but the DEM and tiles.
static int light_args(const lsh_tile* tiles, int32_t tile_count, int32_t dem_width,
                      int32_t dem_height, const float* horizons, int32_t azimuth_count,
                      int32_t azimuth_major, const double* sun, int32_t sun_count,
                      const double* earth, int32_t earth_count, const double* psr_vectors,
                      int32_t psr_count, double sun_half_angle_deg, double psr_diameter_deg,
                      int32_t flags, uint8_t* sun_bytes, float* sun_fraction, float* sun_mean,
                      float* earth_margin, float* earth_visible, uint8_t* psr, float* psr_margin,
                      lsh::LightArgs* out) {
  LSH_REQUIRE(tiles && tile_count > 0 && tile_count <= 65535, "tile_count must be 1..65535");
  LSH_REQUIRE(horizons && azimuth_count >= 2, "horizons and azimuth_count >= 2 are required");
  LSH_REQUIRE(sun_count >= 0 && earth_count >= 0 && psr_count >= 0, "counts must be >= 0");
  LSH_REQUIRE(!sun_count || sun, "sun vectors missing");
  LSH_REQUIRE(!earth_count || earth, "earth vectors missing");
  LSH_REQUIRE(!psr_count || psr_vectors, "PSR vectors missing");
  LSH_REQUIRE(std::isfinite(sun_half_angle_deg) && sun_half_angle_deg > 0.0 &&
                  std::isfinite(psr_diameter_deg) && psr_diameter_deg > 0.0,
              "disk sizes must be positive");
  for (int i = 0; i < tile_count; ++i)
    LSH_REQUIRE(tiles[i].tile_column >= 0 && tiles[i].tile_row >= 0 &&
                    tiles[i].tile_column < dem_width && tiles[i].tile_row < dem_height,
                "tile origin lies outside the DEM");
  LSH_REQUIRE(!(flags & LSH_LIGHT_SUN_BYTES) || sun_bytes, "sun_bytes output missing");
  LSH_REQUIRE(!(flags & LSH_LIGHT_SUN_FRACTION) || sun_fraction, "sun_fraction output missing");
  LSH_REQUIRE(!(flags & LSH_LIGHT_SUN_MEAN) || sun_mean, "sun_mean output missing");
  LSH_REQUIRE(!(flags & LSH_LIGHT_EARTH_MARGIN) || earth_margin, "earth_margin output missing");
  LSH_REQUIRE(!(flags & LSH_LIGHT_EARTH_VISIBLE) || earth_visible, "earth_visible output missing");
  LSH_REQUIRE(!(flags & LSH_LIGHT_PSR) || psr, "psr output missing");
  LSH_REQUIRE(!(flags & LSH_LIGHT_PSR_MARGIN) || psr_margin, "psr_margin output missing");
  lsh::LightArgs& a = *out;
  a = lsh::LightArgs{};
  a.dem_width = dem_width;
  a.dem_height = dem_height;
  a.tile_count = tile_count;
  a.horizons = horizons;
  a.N = azimuth_count;
  a.azimuth_major = azimuth_major ? 1 : 0;
  a.sun = sun;
  a.sun_count = sun_count;
  a.earth = earth;
  a.earth_count = earth_count;
  a.psr_vectors = psr_vectors;
  a.psr_count = psr_count;
  lsh::sun_disk(sun_half_angle_deg, a.half_circle, &a.max_photons);
  a.sun_half_angle = static_cast<float>(sun_half_angle_deg);
  a.psr_threshold = -static_cast<double>(static_cast<float>(psr_diameter_deg)) / 2.0;
  a.sun_bytes = (flags & LSH_LIGHT_SUN_BYTES) ? sun_bytes : nullptr;
  a.sun_fraction = (flags & LSH_LIGHT_SUN_FRACTION) ? sun_fraction : nullptr;
  a.sun_mean = (flags & LSH_LIGHT_SUN_MEAN) ? sun_mean : nullptr;
  a.earth_margin = (flags & LSH_LIGHT_EARTH_MARGIN) ? earth_margin : nullptr;
  a.earth_visible = (flags & LSH_LIGHT_EARTH_VISIBLE) ? earth_visible : nullptr;
  a.psr = (flags & LSH_LIGHT_PSR) ? psr : nullptr;
  a.psr_margin = (flags & LSH_LIGHT_PSR_MARGIN) ? psr_margin : nullptr;
  return LSH_OK;
}

int lsh_lighting_tiles(lsh_session* session, int32_t dem_id, const lsh_tile* tiles,
                       int32_t tile_count, const float* horizons, int32_t azimuth_count,
                       int32_t azimuth_major, const double* sun, int32_t sun_count,
                       const double* earth, int32_t earth_count, const double* psr_vectors,
                       int32_t psr_count, double sun_half_angle_deg, double psr_diameter_deg,
                       int32_t flags, uint8_t* sun_bytes, float* sun_fraction, float* sun_mean,
                       float* earth_margin, float* earth_visible, uint8_t* psr,
                       float* psr_margin) {
  Session* s = reinterpret_cast<Session*>(session);
  Dem* d = find_dem(s, dem_id);
  LSH_REQUIRE(d, "unknown DEM id");
  lsh::LightArgs a;
  if (int r = light_args(tiles, tile_count, d->width, d->height, horizons, azimuth_count,
                         azimuth_major, sun, sun_count, earth, earth_count, psr_vectors, psr_count,
                         sun_half_angle_deg, psr_diameter_deg, flags, sun_bytes, sun_fraction,
                         sun_mean, earth_margin, earth_visible, psr, psr_margin, &a))
    return r;
  a.elevation = d->level0;
  std::memcpy(a.geo_transform, d->geo_transform, sizeof(a.geo_transform));
  std::memcpy(a.projection, d->projection, sizeof(a.projection));
  if (int r = select_device(s)) return r;
  std::lock_guard<std::mutex> lock(s->mutex);
  if (int r = lsh::upload_tiles(s, tiles, tile_count, &a.tiles)) return r;
  LSH_CHECK_HIP(lsh::launch_lighting(a, s->stream), "lighting kernel launch");
  return LSH_OK;
}

// -- lighting without allocation ---------------------------------------------------
// The only device memory lsh_lighting_tiles uses beyond the caller's inputs and
// outputs is the device copy of the tile list (session scratch, grown with
// hipMalloc). lsh_lighting_tiles_static keeps that copy in this file-scope array
// instead: 4096 tiles (32 KiB) per device, held until the process ends. Host
// code never passes `scratch` itself to a kernel: its device address on each
// device comes from hipGetSymbolAddress, taken once per device with that device
// current, and only that address is used.
namespace {

constexpr int32_t kStaticTileCapacity = 4096;
static __device__ __align__(256) lsh_tile scratch[kStaticTileCapacity];

constexpr int kStaticDevices = 64;
std::mutex g_scratch_mutex;
lsh_tile* g_scratch_address[kStaticDevices] = {};

// The device address of `scratch` on the current device (set up on first use).
int static_tile_table(int device, lsh_tile** out) {
  LSH_REQUIRE(device >= 0 && device < kStaticDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_address[device]) {
    void* address = nullptr;
    LSH_CHECK_HIP(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)),
                  "hipGetSymbolAddress(scratch)");
    g_scratch_address[device] = static_cast<lsh_tile*>(address);
  }
  *out = g_scratch_address[device];
  return LSH_OK;
}

}  // namespace

int lsh_lighting_tiles_static(lsh_session* session, int32_t dem_id, const lsh_tile* tiles,
                              int32_t tile_count, const float* horizons, int32_t azimuth_count,
                              int32_t azimuth_major, const double* sun, int32_t sun_count,
                              const double* earth, int32_t earth_count,
                              const double* psr_vectors, int32_t psr_count,
                              double sun_half_angle_deg, double psr_diameter_deg, int32_t flags,
                              uint8_t* sun_bytes, float* sun_fraction, float* sun_mean,
                              float* earth_margin, float* earth_visible, uint8_t* psr,
                              float* psr_margin) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s, "invalid session");
  Dem* d = find_dem(s, dem_id);
  LSH_REQUIRE(d, "unknown DEM id");
  LSH_REQUIRE(tile_count <= kStaticTileCapacity,
              "lsh_lighting_tiles_static takes at most 4096 tiles per call");
  lsh::LightArgs a;
  if (int r = light_args(tiles, tile_count, d->width, d->height, horizons, azimuth_count,
                         azimuth_major, sun, sun_count, earth, earth_count, psr_vectors, psr_count,
                         sun_half_angle_deg, psr_diameter_deg, flags, sun_bytes, sun_fraction,
                         sun_mean, earth_margin, earth_visible, psr, psr_margin, &a))
    return r;
  a.elevation = d->level0;
  std::memcpy(a.geo_transform, d->geo_transform, sizeof(a.geo_transform));
  std::memcpy(a.projection, d->projection, sizeof(a.projection));
  if (int r = select_device(s)) return r;
  lsh_tile* table = nullptr;
  if (int r = static_tile_table(s->device, &table)) return r;
  std::lock_guard<std::mutex> lock(s->mutex);
  // stream ordered: a previous call's kernel on this stream has read the table
  // before this copy replaces it (calls on other streams of the device race)
  LSH_CHECK_HIP(hipMemcpyAsync(table, tiles, sizeof(lsh_tile) * tile_count,
                               hipMemcpyHostToDevice, s->stream),
                "copy tile list");
  a.tiles = table;
  LSH_CHECK_HIP(lsh::launch_lighting(a, s->stream), "lighting kernel launch");
  return LSH_OK;
}

int lsh_sun_loss(lsh_session* session, int32_t dem_id, const lsh_tile* tiles,
                 int32_t tile_count, const float* horizons, int32_t azimuth_count,
                 const double* sun, int32_t sun_count, double sun_half_angle_deg,
                 double smoothing_deg, double* loss, float* grad_horizon) {
  Session* s = reinterpret_cast<Session*>(session);
  Dem* d = find_dem(s, dem_id);
  LSH_REQUIRE(d, "unknown DEM id");
  LSH_REQUIRE(loss || grad_horizon, "nothing to compute");
  LSH_REQUIRE(sun_count >= 1, "sun vectors are required");
  // the gradient keeps the <= 5 bins one time's slice walk touches at N <= 1440
  LSH_REQUIRE(azimuth_count <= 1440, "the sun loss takes at most 1440 azimuths");
  lsh::LightArgs a;
  if (int r = light_args(tiles, tile_count, d->width, d->height, horizons, azimuth_count, 0, sun,
                         sun_count, nullptr, 0, nullptr, 0, sun_half_angle_deg, 0.545, 0,
                         nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &a))
    return r;
  LSH_REQUIRE(std::isfinite(smoothing_deg) && smoothing_deg > 0.0 &&
                  2.0 * smoothing_deg <= 2.0 * a.half_circle[0],
              "smoothing must be positive and at most the narrowest slice half-height");
  a.elevation = d->level0;
  std::memcpy(a.geo_transform, d->geo_transform, sizeof(a.geo_transform));
  std::memcpy(a.projection, d->projection, sizeof(a.projection));
  if (int r = select_device(s)) return r;
  std::lock_guard<std::mutex> lock(s->mutex);
  if (int r = lsh::upload_tiles(s, tiles, tile_count, &a.tiles)) return r;
  LSH_CHECK_HIP(lsh::launch_sun_loss(a, smoothing_deg, loss, grad_horizon, s->stream),
                "sun loss kernel launch");
  return LSH_OK;
}

int lsh_lighting_tiles_host(const float* elevation, int32_t width, int32_t height,
                            const double* geo_transform, const double* projection,
                            const lsh_tile* tiles, int32_t tile_count, const float* horizons,
                            int32_t azimuth_count, int32_t azimuth_major, const double* sun,
                            int32_t sun_count, const double* earth, int32_t earth_count,
                            const double* psr_vectors, int32_t psr_count,
                            double sun_half_angle_deg, double psr_diameter_deg, int32_t flags,
                            int32_t threads, uint8_t* sun_bytes, float* sun_fraction,
                            float* sun_mean, float* earth_margin, float* earth_visible,
                            uint8_t* psr, float* psr_margin) {
  LSH_REQUIRE(elevation && geo_transform && projection && width > 0 && height > 0,
              "DEM arguments are invalid");
  lsh::LightArgs a;
  if (int r = light_args(tiles, tile_count, width, height, horizons, azimuth_count, azimuth_major,
                         sun, sun_count, earth, earth_count, psr_vectors, psr_count,
                         sun_half_angle_deg, psr_diameter_deg, flags, sun_bytes, sun_fraction,
                         sun_mean, earth_margin, earth_visible, psr, psr_margin, &a))
    return r;
  a.elevation = elevation;
  std::memcpy(a.geo_transform, geo_transform, sizeof(a.geo_transform));
  std::memcpy(a.projection, projection, sizeof(a.projection));
  a.tiles = tiles;
  lsh::run_lighting_host(a, threads);
  return LSH_OK;
}

int lsh_psr_reduce(const float* elevation, int32_t width, int32_t height,
                   const double* geo_transform, const double* projection, const double* vectors,
                   int32_t count, int32_t azimuth_count, int64_t* indices, int32_t* kept) {
  LSH_REQUIRE(elevation && geo_transform && projection && indices && kept, "invalid arguments");
  LSH_REQUIRE(width > 0 && height > 0 && azimuth_count > 0, "invalid dimensions");
  LSH_REQUIRE(count > 0 && vectors, "vectors must be non-empty");
  for (int64_t i = 0; i < 3 * static_cast<int64_t>(count); ++i)
    LSH_REQUIRE(std::isfinite(vectors[i]), "vectors must be finite");
  const double kPi = 3.14159265358979323846;
  const int points[5][2] = {{0, 0}, {0, width - 1}, {height - 1, 0}, {height - 1, width - 1},
                            {height / 2, width / 2}};
  std::vector<int64_t> selected(5 * static_cast<size_t>(azimuth_count), -1);
  std::vector<double> best(5 * static_cast<size_t>(azimuth_count), -HUGE_VAL);
  for (int k = 0; k < 5; ++k) {
    int row = points[k][0], column = points[k][1];
    const double* t = geo_transform;
    const double* p = projection;
    double x = t[0] + t[1] * column + t[2] * row, y = t[3] + t[4] * column + t[5] * row;
    double xp = x - p[4], yp = y - p[5], rho = std::hypot(xp, yp), lat, lon;
    if (rho < 1e-9) {
      lat = p[1];
      lon = p[2];
    } else {
      double c = 2.0 * std::atan2(rho, 2.0 * p[3] * p[0]);
      lat = std::asin(std::cos(c) * std::sin(p[1]) + yp * std::sin(c) * std::cos(p[1]) / rho);
      lon = p[2] + std::atan2(xp * std::sin(c),
                              rho * std::cos(p[1]) * std::cos(c) - yp * std::sin(p[1]) * std::sin(c));
    }
    double cl = std::cos(lat), sl = std::sin(lat), co = std::cos(lon), so = std::sin(lon);
    double basis[3][3] = {{-so, co, 0.0}, {-sl * co, -sl * so, cl}, {cl * co, cl * so, sl}};
    double radius = p[0] + static_cast<double>(elevation[static_cast<int64_t>(row) * width + column]);
    double obs[3] = {radius * basis[2][0], radius * basis[2][1], radius * basis[2][2]};
    double tr[3];
    for (int j = 0; j < 3; ++j)
      tr[j] = -(obs[0] * basis[j][0] + obs[1] * basis[j][1] + obs[2] * basis[j][2]);
    for (int i = 0; i < count; ++i) {
      const double* v = vectors + 3 * static_cast<int64_t>(i);
      double e[3];
      for (int j = 0; j < 3; ++j) e[j] = v[0] * basis[j][0] + v[1] * basis[j][1] + v[2] * basis[j][2] + tr[j];
      double az = std::fmod(std::atan2(e[0], e[1]) * (180.0 / kPi), 360.0);
      if (az < 0.0) az += 360.0;
      double el = std::atan2(e[2], std::hypot(e[0], e[1])) * (180.0 / kPi);
      int64_t bin = static_cast<int64_t>(az * (azimuth_count / 360.0)) % azimuth_count;
      size_t idx = static_cast<size_t>(k) * azimuth_count + bin;
      if (el > best[idx]) {
        best[idx] = el;
        selected[idx] = i;
      }
    }
  }
  std::vector<char> seen(count, 0);
  int32_t n = 0;
  for (int64_t v : selected)
    if (v >= 0 && !seen[v]) {
      seen[v] = 1;
      indices[n++] = v;
    }
  *kept = n;
  return LSH_OK;
}

int lsh_probe_copy(lsh_session* session, void* device_dst, const void* device_src, size_t bytes) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && device_dst && device_src, "invalid arguments");
  LSH_REQUIRE(bytes % 16 == 0, "bytes must be a multiple of 16");
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(lsh::launch_probe_copy(static_cast<const float*>(device_src),
                                  static_cast<float*>(device_dst), static_cast<int64_t>(bytes),
                                  s->stream),
                "copy probe launch");
  return LSH_OK;
}

int lsh_probe_gather(lsh_session* session, const float* window, int32_t width, int32_t height,
                     int32_t taps, int32_t threads, float* out) {
  Session* s = reinterpret_cast<Session*>(session);
  LSH_REQUIRE(s && window && out, "invalid arguments");
  LSH_REQUIRE(width >= 2 && height >= 2 && width <= (1 << 24) && height <= (1 << 24),
              "window must be at least 2x2 and at most 2^24 on a side");
  LSH_REQUIRE(taps >= 1 && threads >= 256 && threads % 256 == 0,
              "taps must be positive and threads a positive multiple of 256");
  if (int r = select_device(s)) return r;
  LSH_CHECK_HIP(lsh::launch_probe_gather(window, width, height, taps, threads, out, s->stream),
                "gather probe launch");
  return LSH_OK;
}

}  // extern "C"
