// lsh_durations.hip -- the longest-duration and safe-haven reductions of the
// lighting time series (lsh_durations_core.h), on the GPU and on host threads.
//
// One thread per pixel, as the lighting kernel: every thread of a workgroup
// reads the same vector and the same spec entry (uniform scalar loads) and its
// own horizon row; the per-band state lives in the workspace at
// [band][tile][pixel], so a wavefront's state accesses are contiguous, and is
// touched only when a run ends or the month changes.

#include <cstring>
#include <thread>
#include <vector>

#include "lsh_durations_core.h"

namespace lsh {
namespace {

template <bool kAzimuthMajor>
__global__ void __launch_bounds__(256) duration_kernel(DurationArgs a) {
  // each thread's sector tables of the sun signal and the Earth margin,
  // [sector][thread]
  __shared__ float s_dark[kSectors][256], s_lit[kSectors][256];
  __shared__ float s_earth_dark[kSectors / 2][256], s_earth_lit[kSectors / 2][256];
  const int tile = blockIdx.y;
  const int pixel = blockIdx.x * blockDim.x + threadIdx.x;
  if (pixel >= kPatchPixels) return;
  const SectorTable sectors{&s_dark[0][threadIdx.x], &s_lit[0][threadIdx.x], 256};
  const SectorTable earth_sectors{&s_earth_dark[0][threadIdx.x], &s_earth_lit[0][threadIdx.x], 256};
  duration_pixel<kAzimuthMajor>(a, tile, pixel, a.light.tiles[tile], sectors, earth_sectors);
}

}  // namespace

hipError_t launch_durations(const DurationArgs& a, hipStream_t stream) {
  dim3 block(256), grid(kPatchPixels / 256, a.light.tile_count);
  if (a.light.azimuth_major)
    hipLaunchKernelGGL(duration_kernel<true>, grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL(duration_kernel<false>, grid, block, 0, stream, a);
  return hipGetLastError();
}

void run_durations_host(const DurationArgs& a, int threads) {
  const int64_t items = static_cast<int64_t>(a.light.tile_count) * kPatchPixels;
  int n = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
  n = n < 1 ? 1 : (n > 256 ? 256 : n);
  auto body = [&a, items, n](int t) {
    const int64_t begin = items * t / n, end = items * (t + 1) / n;
    float dark[kSectors], lit[kSectors], earth_dark[kSectors / 2], earth_lit[kSectors / 2];
    const SectorTable sectors{dark, lit, 1}, earth_sectors{earth_dark, earth_lit, 1};
    for (int64_t i = begin; i < end; ++i) {
      const int tile = static_cast<int>(i / kPatchPixels), pixel = static_cast<int>(i % kPatchPixels);
      if (a.light.azimuth_major)
        duration_pixel<true>(a, tile, pixel, a.light.tiles[tile], sectors, earth_sectors);
      else
        duration_pixel<false>(a, tile, pixel, a.light.tiles[tile], sectors, earth_sectors);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < n; ++t) pool.emplace_back(body, t);
  body(0);
  for (auto& th : pool) th.join();
}

// Workspace layout: longest double [K][tiles][16384] (longest duration), or
// best int32 [B][tiles][16384] followed by flags uint8 [B][tiles][16384].
int64_t duration_workspace_bytes(const lsh_duration_spec& spec, int32_t tile_count) {
  const int64_t cells = static_cast<int64_t>(tile_count) * kPatchPixels;
  if (spec.reducer == LSH_REDUCE_SAFE_HAVEN)
    return cells * spec.band_count * static_cast<int64_t>(sizeof(int32_t) + sizeof(uint8_t));
  return cells * spec.interval_count * static_cast<int64_t>(sizeof(double));
}

void bind_workspace(DurationArgs& a, void* workspace) {
  const int64_t cells = static_cast<int64_t>(a.light.tile_count) * kPatchPixels;
  if (a.reducer == LSH_REDUCE_SAFE_HAVEN) {
    a.best = static_cast<int32_t*>(workspace);
    a.flags = reinterpret_cast<uint8_t*>(a.best + cells * a.B);
  } else {
    a.longest = static_cast<double*>(workspace);
  }
}

}  // namespace lsh

// ---------------------------------------------------------------------------
// C ABI (declared in include/lunarscout_hip.h)

namespace {

// Validates the spec and the shared arguments and fills everything but the
// DEM fields, the tile list, the spec tables and the workspace.
int duration_args(const lsh_tile* tiles, int32_t tile_count, int32_t dem_width,
                  int32_t dem_height, const float* horizons, int32_t azimuth_count,
                  int32_t azimuth_major, const double* sun, const double* earth,
                  double sun_half_angle_deg, const lsh_duration_spec* spec, void* workspace,
                  float* out, lsh::DurationArgs* a) {
  LSH_REQUIRE(spec && workspace && out, "spec, workspace and out are required");
  LSH_REQUIRE(tiles && tile_count > 0 && tile_count <= 65535, "tile_count must be 1..65535");
  LSH_REQUIRE(horizons && azimuth_count >= 2, "horizons and azimuth_count >= 2 are required");
  LSH_REQUIRE(std::isfinite(sun_half_angle_deg) && sun_half_angle_deg > 0.0,
              "the sun half angle must be positive");
  for (int i = 0; i < tile_count; ++i)
    LSH_REQUIRE(tiles[i].tile_column >= 0 && tiles[i].tile_row >= 0 &&
                    tiles[i].tile_column < dem_width && tiles[i].tile_row < dem_height,
                "tile origin lies outside the DEM");
  const lsh_duration_spec& p = *spec;
  const int T = p.time_count;
  LSH_REQUIRE(T >= 1 && sun, "time_count >= 1 and sun vectors are required");
  LSH_REQUIRE(p.reducer == LSH_REDUCE_LONGEST_DURATION || p.reducer == LSH_REDUCE_SAFE_HAVEN,
              "unknown reducer");
  LSH_REQUIRE(std::isfinite(p.sun_threshold), "the sun threshold must be finite");
  const bool need_earth = p.reducer == LSH_REDUCE_SAFE_HAVEN || p.use_earth;
  LSH_REQUIRE(!need_earth || (earth && std::isfinite(p.earth_threshold_deg)),
              "earth vectors and a finite earth threshold are required");
  if (p.reducer == LSH_REDUCE_LONGEST_DURATION) {
    LSH_REQUIRE(p.sun_signal == LSH_SIGNAL_SUN_FRACTION || p.sun_signal == LSH_SIGNAL_SUN_MARGIN,
                "unknown sun signal");
    LSH_REQUIRE(p.sample_hours && p.interval_count >= 1 && p.interval_first && p.interval_end,
                "sample hours and at least one interval are required");
    for (int t = 0; t < T; ++t)
      LSH_REQUIRE(std::isfinite(p.sample_hours[t]) && p.sample_hours[t] >= 0.0,
                  "sample hours must be finite and non-negative");
    for (int k = 0; k < p.interval_count; ++k)
      LSH_REQUIRE(p.interval_first[k] >= 0 && p.interval_first[k] <= p.interval_end[k] &&
                      p.interval_end[k] <= T,
                  "interval sample ranges must lie in [0, time_count]");
  } else {
    LSH_REQUIRE(p.month_of && p.band_count >= 1, "month_of and band_count >= 1 are required");
    LSH_REQUIRE(std::isfinite(p.step_hours) && p.step_hours > 0.0, "step_hours must be positive");
    for (int t = 0; t < T; ++t)
      LSH_REQUIRE(p.month_of[t] >= -1 && p.month_of[t] < p.band_count,
                  "month_of entries must lie in [-1, band_count)");
  }
  bool sorted = p.reducer == LSH_REDUCE_SAFE_HAVEN;
  for (int t = 1; sorted && t < T; ++t) sorted = p.month_of[t] >= p.month_of[t - 1];
  lsh::DurationArgs& d = *a;
  d = lsh::DurationArgs{};
  d.light.dem_width = dem_width;
  d.light.dem_height = dem_height;
  d.light.tile_count = tile_count;
  d.light.horizons = horizons;
  d.light.N = azimuth_count;
  d.light.azimuth_major = azimuth_major ? 1 : 0;
  d.light.sun = sun;
  d.light.sun_count = T;
  d.light.earth = need_earth ? earth : nullptr;
  d.light.earth_count = need_earth ? T : 0;
  lsh::sun_disk(sun_half_angle_deg, d.light.half_circle, &d.light.max_photons);
  d.light.sun_half_angle = static_cast<float>(sun_half_angle_deg);
  d.reducer = p.reducer;
  d.sun_signal = p.reducer == LSH_REDUCE_SAFE_HAVEN ? LSH_SIGNAL_SUN_FRACTION : p.sun_signal;
  d.use_earth = need_earth ? 1 : 0;
  d.T = T;
  d.sun_threshold = static_cast<float>(p.sun_threshold);
  d.earth_threshold = static_cast<float>(p.earth_threshold_deg);
  d.K = p.reducer == LSH_REDUCE_LONGEST_DURATION ? p.interval_count : 0;
  d.output_days = p.output_days ? 1 : 0;
  d.B = p.reducer == LSH_REDUCE_SAFE_HAVEN ? p.band_count : 0;
  d.months_sorted = sorted ? 1 : 0;
  d.step_hours = static_cast<float>(p.step_hours);
  d.out = out;
  return LSH_OK;
}

// The spec's tables packed as one block: sample hours, interval ranges, months.
std::vector<unsigned char> pack_tables(const lsh_duration_spec& p, size_t* hours_at,
                                       size_t* first_at, size_t* end_at, size_t* month_at) {
  const size_t T = static_cast<size_t>(p.time_count);
  const size_t K = p.reducer == LSH_REDUCE_LONGEST_DURATION ? p.interval_count : 0;
  const bool months = p.reducer == LSH_REDUCE_SAFE_HAVEN;
  *hours_at = 0;
  const size_t hours_bytes = K ? sizeof(double) * T : 0;
  *first_at = hours_bytes;
  *end_at = *first_at + sizeof(int32_t) * K;
  *month_at = *end_at + sizeof(int32_t) * K;
  std::vector<unsigned char> block(*month_at + (months ? sizeof(int32_t) * T : 0) + 8);
  if (K) {
    std::memcpy(block.data() + *hours_at, p.sample_hours, hours_bytes);
    std::memcpy(block.data() + *first_at, p.interval_first, sizeof(int32_t) * K);
    std::memcpy(block.data() + *end_at, p.interval_end, sizeof(int32_t) * K);
  }
  if (months) std::memcpy(block.data() + *month_at, p.month_of, sizeof(int32_t) * T);
  return block;
}

}  // namespace

extern "C" {

int lsh_duration_workspace_bytes(const lsh_duration_spec* spec, int32_t tile_count,
                                 int64_t* bytes) {
  LSH_REQUIRE(spec && bytes && tile_count > 0, "invalid arguments");
  *bytes = lsh::duration_workspace_bytes(*spec, tile_count);
  return LSH_OK;
}

int lsh_duration_tiles(lsh_session* session, int32_t dem_id, const lsh_tile* tiles,
                       int32_t tile_count, const float* horizons, int32_t azimuth_count,
                       int32_t azimuth_major, const double* sun, const double* earth,
                       double sun_half_angle_deg, const lsh_duration_spec* spec,
                       void* workspace, float* out) {
  lsh::Session* s = reinterpret_cast<lsh::Session*>(session);
  lsh::Dem* dem = lsh::find_dem(s, dem_id);
  LSH_REQUIRE(dem, "unknown DEM id");
  lsh::DurationArgs a;
  if (int r = duration_args(tiles, tile_count, dem->width, dem->height, horizons, azimuth_count,
                            azimuth_major, sun, earth, sun_half_angle_deg, spec, workspace, out,
                            &a))
    return r;
  a.light.elevation = dem->level0;
  std::memcpy(a.light.geo_transform, dem->geo_transform, sizeof(a.light.geo_transform));
  std::memcpy(a.light.projection, dem->projection, sizeof(a.light.projection));
  size_t hours_at, first_at, end_at, month_at;
  const std::vector<unsigned char> block = pack_tables(*spec, &hours_at, &first_at, &end_at, &month_at);
  if (int r = lsh::select_device(s)) return r;
  std::lock_guard<std::mutex> lock(s->mutex);
  if (int r = lsh::upload_tiles(s, tiles, tile_count, &a.light.tiles)) return r;
  void* tables = nullptr;
  if (int r = lsh::upload_tables(s, block.data(), block.size(), &tables)) return r;
  unsigned char* base = static_cast<unsigned char*>(tables);
  a.sample_hours = reinterpret_cast<const double*>(base + hours_at);
  a.first = reinterpret_cast<const int32_t*>(base + first_at);
  a.end = reinterpret_cast<const int32_t*>(base + end_at);
  a.month_of = reinterpret_cast<const int32_t*>(base + month_at);
  lsh::bind_workspace(a, workspace);
  LSH_CHECK_HIP(lsh::launch_durations(a, s->stream), "duration kernel launch");
  // the table block is host memory owned by this call: wait for the copy
  LSH_CHECK_HIP(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  return LSH_OK;
}

int lsh_duration_tiles_host(const float* elevation, int32_t width, int32_t height,
                            const double* geo_transform, const double* projection,
                            const lsh_tile* tiles, int32_t tile_count, const float* horizons,
                            int32_t azimuth_count, int32_t azimuth_major, const double* sun,
                            const double* earth, double sun_half_angle_deg,
                            const lsh_duration_spec* spec, int32_t threads, void* workspace,
                            float* out) {
  LSH_REQUIRE(elevation && geo_transform && projection && width > 0 && height > 0,
              "DEM arguments are invalid");
  lsh::DurationArgs a;
  if (int r = duration_args(tiles, tile_count, width, height, horizons, azimuth_count,
                            azimuth_major, sun, earth, sun_half_angle_deg, spec, workspace, out,
                            &a))
    return r;
  a.light.elevation = elevation;
  std::memcpy(a.light.geo_transform, geo_transform, sizeof(a.light.geo_transform));
  std::memcpy(a.light.projection, projection, sizeof(a.light.projection));
  a.light.tiles = tiles;
  a.sample_hours = spec->sample_hours;
  a.first = spec->interval_first;
  a.end = spec->interval_end;
  a.month_of = spec->month_of;
  lsh::bind_workspace(a, workspace);
  lsh::run_durations_host(a, threads);
  return LSH_OK;
}

}  // extern "C"
