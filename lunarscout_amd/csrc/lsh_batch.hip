// lsh_batch.hip -- batched horizon generation: many tiles of many sites per launch.
//
// A batch holds, for its sites, the device site tables of the segment and march
// kernels, a segment slot per tile, a pinned staging copy of the tile list, the
// eight work-queue counters and the hipGraph of the per-batch sequence:
//   1. copy the tile list from pinned staging to the device table,
//   2. FP64 segments of every tile (one launch, blockIdx.y = tile),
//   3. reset the queue counters (a fill kernel: no memset node in the graph),
//   4. the persistent march (march_batch_kernel) over all tiles.
// Each run rewrites the staging list and replays the graph; the graph is
// captured again only when the tile count, kernel shape, workgroup count,
// flags or output pointer change. The per-ray arithmetic is the plain path's.

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lsh_internal.h"

struct lsh_batch {
  lsh::Session* s = nullptr;
  int32_t max_tiles = 0, N = 0, C = 0, count = 0, subpatch = 8, max_dems = 1;
  double maximum_distance = 1e6;
  std::vector<int32_t> site_width, site_height;
  lsh::MarchSite* d_march_sites = nullptr;
  lsh::SegmentSite* d_segment_sites = nullptr;
  lsh_batch_tile* h_tiles = nullptr;  // pinned staging
  lsh_batch_tile* d_tiles = nullptr;
  float* d_segments = nullptr;
  int64_t slot_floats = 0;
  unsigned int* d_queues = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int32_t g_tiles = -1, g_config = -1, g_workgroups = -1, g_flags = -1;
  float* g_output = nullptr;
  hipEvent_t done = nullptr;
  bool pending = false;
};

namespace {

void release(lsh_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->done) hipEventDestroy(b->done);
  if (b->d_march_sites) hipFree(b->d_march_sites);
  if (b->d_segment_sites) hipFree(b->d_segment_sites);
  if (b->h_tiles) hipHostFree(b->h_tiles);
  if (b->d_tiles) hipFree(b->d_tiles);
  if (b->d_segments) hipFree(b->d_segments);
  if (b->d_queues) hipFree(b->d_queues);
  delete b;
}

// The per-batch sequence on the session stream (captured or direct).
hipError_t enqueue(lsh_batch* b, int32_t tile_count, int32_t config, int32_t workgroups,
                   int32_t flags, float* output) {
  hipStream_t stream = b->s->stream;
  hipError_t e = hipMemcpyAsync(b->d_tiles, b->h_tiles, sizeof(lsh_batch_tile) * tile_count,
                                hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return e;
  lsh::SegmentBatchArgs sa{};
  sa.sites = b->d_segment_sites;
  sa.tiles = b->d_tiles;
  sa.segments = b->d_segments;
  sa.slot_floats = b->slot_floats;
  sa.tile_count = tile_count;
  sa.N = b->N;
  sa.C = b->C;
  sa.count = b->count;
  sa.subpatch = b->subpatch;
  sa.maximum_distance = b->maximum_distance;
  if ((e = lsh::launch_segments_batch(sa, stream)) != hipSuccess) return e;
  if ((e = lsh::launch_reset_queues(b->d_queues, stream)) != hipSuccess) return e;
  lsh::BatchMarchArgs ma{};
  ma.sites = b->d_march_sites;
  ma.tiles = b->d_tiles;
  ma.segments = b->d_segments;
  ma.slot_floats = b->slot_floats;
  ma.output = output;
  ma.queues = b->d_queues;
  ma.tile_count = tile_count;
  ma.N = b->N;
  ma.C = b->C;
  ma.subpatch = b->subpatch;
  ma.flags = flags;
  return lsh::launch_march_batch(ma, config, workgroups, stream);
}

}  // namespace

extern "C" {

int lsh_batch_shape_count(int32_t* count) {
  LSH_REQUIRE(count, "count must not be null");
  *count = lsh::batch_shape_count();
  return LSH_OK;
}

int lsh_batch_shape(int32_t config, int32_t* block_x, int32_t* block_y, int32_t* azimuths) {
  LSH_REQUIRE(block_x && block_y && azimuths, "invalid arguments");
  LSH_REQUIRE(config >= 0 && config < lsh::batch_shape_count(), "unknown kernel shape");
  const lsh::BatchShape shape = lsh::batch_shape(config);
  *block_x = shape.block_x;
  *block_y = shape.block_y;
  *azimuths = shape.azimuths;
  return LSH_OK;
}

int lsh_batch_resident_workgroups(lsh_session* session, int32_t config, int32_t* workgroups) {
  lsh::Session* s = reinterpret_cast<lsh::Session*>(session);
  LSH_REQUIRE(s && workgroups, "invalid arguments");
  LSH_REQUIRE(config >= 0 && config < lsh::batch_shape_count(), "unknown kernel shape");
  if (int r = lsh::select_device(s)) return r;
  *workgroups = lsh::batch_default_workgroups(config, s->cu_count);
  return LSH_OK;
}

int lsh_batch_create(lsh_session* session, const lsh_batch_site* sites, int32_t site_count,
                     int32_t max_tiles, int32_t azimuth_count, double maximum_distance_m,
                     lsh_batch** batch) {
  lsh::Session* s = reinterpret_cast<lsh::Session*>(session);
  LSH_REQUIRE(s && sites && batch, "invalid arguments");
  LSH_REQUIRE(site_count >= 1 && max_tiles >= 1 && max_tiles <= 65535, "invalid batch size");
  LSH_REQUIRE(azimuth_count >= 1, "azimuth_count must be positive");
  LSH_REQUIRE(std::isfinite(maximum_distance_m) && maximum_distance_m > 0.0,
              "maximum distance must be positive");
  std::vector<lsh::MarchSite> march(site_count);
  std::vector<lsh::SegmentSite> segment(site_count);
  lsh_batch* b = new lsh_batch();
  b->s = s;
  b->max_tiles = max_tiles;
  b->N = azimuth_count;
  b->count = 128 / b->subpatch + 2;
  b->C = b->count * b->count;
  b->maximum_distance = maximum_distance_m;
  for (int i = 0; i < site_count; ++i) {
    const lsh_batch_site& site = sites[i];
    if (!(site.dem_count >= 1 && site.dem_count <= lsh::kMaxDems) ||
        !std::isfinite(site.observer_elevation_m)) {
      delete b;
      lsh::set_error("each site needs 1..4 DEMs and a finite observer height");
      return LSH_EINVAL;
    }
    march[i] = lsh::MarchSite{};
    segment[i] = lsh::SegmentSite{};
    march[i].D = segment[i].D = site.dem_count;
    march[i].observer_elevation = site.observer_elevation_m;
    segment[i].observer_elevation = site.observer_elevation_m;
    for (int d = 0; d < site.dem_count; ++d) {
      lsh::Dem* dem = lsh::find_dem(s, site.dem_ids[d]);
      if (!dem) {
        delete b;
        lsh::set_error("unknown DEM id in a batch site");
        return LSH_EINVAL;
      }
      march[i].dem[d] = lsh::march_dem(*dem);
      segment[i].dem[d] = lsh::segment_dem(*dem, maximum_distance_m);
    }
    b->max_dems = site.dem_count > b->max_dems ? site.dem_count : b->max_dems;
    b->site_width.push_back(march[i].dem[0].width);
    b->site_height.push_back(march[i].dem[0].height);
  }
  b->slot_floats = static_cast<int64_t>(b->N) * b->C * b->max_dems * lsh::kSegmentFields;
  hipError_t e = hipSetDevice(s->device);
  if (e == hipSuccess) e = hipMalloc(&b->d_march_sites, sizeof(lsh::MarchSite) * site_count);
  if (e == hipSuccess) e = hipMalloc(&b->d_segment_sites, sizeof(lsh::SegmentSite) * site_count);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_tiles, sizeof(lsh_batch_tile) * max_tiles);
  if (e == hipSuccess) e = hipMalloc(&b->d_tiles, sizeof(lsh_batch_tile) * max_tiles);
  if (e == hipSuccess)
    e = hipMalloc(&b->d_segments, sizeof(float) * b->slot_floats * max_tiles);
  if (e == hipSuccess) e = hipMalloc(&b->d_queues, sizeof(unsigned int) * 8);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_march_sites, march.data(), sizeof(lsh::MarchSite) * site_count,
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_segment_sites, segment.data(), sizeof(lsh::SegmentSite) * site_count,
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    lsh::set_error(std::string("batch allocation: ") + hipGetErrorString(e));
    release(b);
    return e == hipErrorOutOfMemory ? LSH_ENOMEM : LSH_EDEVICE;
  }
  *batch = b;
  return LSH_OK;
}

int lsh_batch_run(lsh_batch* b, const lsh_batch_tile* tiles, int32_t tile_count, int32_t config,
                  int32_t workgroups, int32_t flags, int32_t use_graph, float* output) {
  LSH_REQUIRE(b && tiles && output, "invalid arguments");
  LSH_REQUIRE(tile_count >= 1 && tile_count <= b->max_tiles, "tile_count exceeds the batch size");
  LSH_REQUIRE(config >= 0 && config < lsh::batch_shape_count(), "unknown kernel shape");
  LSH_REQUIRE((flags & ~(LSH_HORIZON_DEGREES | LSH_HORIZON_AZIMUTH_MAJOR)) == 0,
              "batch flags: LSH_HORIZON_DEGREES and/or LSH_HORIZON_AZIMUTH_MAJOR");
  LSH_REQUIRE(workgroups <= 0 || workgroups % 8 == 0, "workgroups must be a multiple of 8");
  for (int i = 0; i < tile_count; ++i) {
    const lsh_batch_tile& t = tiles[i];
    LSH_REQUIRE(t.site >= 0 && t.site < static_cast<int32_t>(b->site_width.size()),
                "tile names an unknown site");
    LSH_REQUIRE(t.tile_column >= 0 && t.tile_row >= 0 && t.tile_column < b->site_width[t.site] &&
                    t.tile_row < b->site_height[t.site] && t.reserved == 0,
                "tile origin lies outside its site's primary DEM");
  }
  lsh::Session* s = b->s;
  if (int r = lsh::select_device(s)) return r;
  std::lock_guard<std::mutex> lock(s->mutex);
  const int wgs = workgroups > 0 ? workgroups : lsh::batch_default_workgroups(config, s->cu_count);
  // the previous run has consumed the staging list once its work is done
  if (b->pending) LSH_CHECK_HIP(hipEventSynchronize(b->done), "previous batch");
  std::memcpy(b->h_tiles, tiles, sizeof(lsh_batch_tile) * tile_count);
  if (use_graph) {
    if (!b->exec || b->g_tiles != tile_count || b->g_config != config ||
        b->g_workgroups != wgs || b->g_flags != flags || b->g_output != output) {
      if (b->exec) hipGraphExecDestroy(b->exec);
      if (b->graph) hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      LSH_CHECK_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal),
                    "begin graph capture");
      hipError_t e = enqueue(b, tile_count, config, wgs, flags, output);
      hipGraph_t graph = nullptr;
      hipError_t end = hipStreamEndCapture(s->stream, &graph);
      LSH_CHECK_HIP(e, "batch capture");
      LSH_CHECK_HIP(end, "end graph capture");
      b->graph = graph;
      LSH_CHECK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0),
                    "graph instantiate");
      b->g_tiles = tile_count;
      b->g_config = config;
      b->g_workgroups = wgs;
      b->g_flags = flags;
      b->g_output = output;
    }
    LSH_CHECK_HIP(hipGraphLaunch(b->exec, s->stream), "graph launch");
  } else {
    LSH_CHECK_HIP(enqueue(b, tile_count, config, wgs, flags, output), "batch launch");
  }
  LSH_CHECK_HIP(hipEventRecord(b->done, s->stream), "batch event");
  b->pending = true;
  return LSH_OK;
}

int lsh_batch_destroy(lsh_batch* b) {
  if (!b) return LSH_OK;
  hipSetDevice(b->s->device);
  if (b->pending) hipEventSynchronize(b->done);
  release(b);
  return LSH_OK;
}

}  // extern "C"
