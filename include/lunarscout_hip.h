/*
 * lunarscout_hip.h -- C ABI of the MI355X horizon and lighting library
 * (liblunarscout_hip.so, built for gfx950).
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types. Every entry
 * point returns LSH_OK (0) or a negative status; lsh_last_error() gives the
 * calling thread's last message. Output buffers are not written when argument
 * validation fails. Buffers marked "device" are device memory obtained from
 * lsh_alloc (or any HIP device allocation on the session's device); the library
 * never frees caller memory.
 *
 * Each entry point names the reference interface it replaces
 * (paths under src/lunarscout/_numba_horizon/ of the reference repository).
 */
#ifndef LUNARSCOUT_HIP_H
#define LUNARSCOUT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. LSH_EINVAL maps to ValueError (cuda_backend.py:1459-1465);
 * LSH_ENODEV to the backend-unavailable error (CudaBackendError,
 * cuda_backend.py:1164-1176); LSH_EDEVICE to an execution failure
 * (horizon.py:426-431); LSH_EIO to the OSError a failed file write raises
 * (file_format.py:398-415). */
#define LSH_OK 0
#define LSH_EINVAL -1
#define LSH_ENODEV -2
#define LSH_EDEVICE -3
#define LSH_ENOMEM -4
#define LSH_EIO -5

/* lsh_horizon_tile / lsh_horizon_batch flags */
#define LSH_HORIZON_DEGREES 1   /* write float32 degrees (contract.py:362-371) instead of slopes */
#define LSH_HORIZON_CAP 2       /* diagnostic: cap level-0 advances at the cell exit (hierarchy.py:258-262) */
#define LSH_HORIZON_COUNTERS 4  /* accumulate traversal counters (hierarchy.py:40-52) */
#define LSH_HORIZON_F32_VALUES 8 /* heights, bounds and slopes in float32 (faster; float32 ties
                                    let rare rays drift from the float64 method) */
#define LSH_HORIZON_AZIMUTH_MAJOR 16 /* output [N][tile_height*tile_width] instead of the
                                        pixel-major tile file order (what the lighting kernel
                                        reads coalesced; lsh_transpose_tiles converts) */

/* lsh_lighting_tiles flags */
#define LSH_LIGHT_SUN_BYTES 1     /* per-time uint8 sun fraction bytes */
#define LSH_LIGHT_SUN_FRACTION 2  /* per-time float32 sun fraction */
#define LSH_LIGHT_SUN_MEAN 4      /* per-pixel mean sun fraction over all times */
#define LSH_LIGHT_EARTH_MARGIN 8  /* per-time float32 Earth margin */
#define LSH_LIGHT_EARTH_VISIBLE 16/* per-pixel fraction of times with Earth margin >= 0 */
#define LSH_LIGHT_PSR 32          /* per-pixel PSR mask (255 = PSR) */
#define LSH_LIGHT_PSR_MARGIN 64   /* per-pixel float32 max(elevation - horizon) over PSR vectors */

typedef struct lsh_session lsh_session;

/* A horizon tile origin within the primary DEM. */
typedef struct {
  int32_t tile_column;
  int32_t tile_row;
} lsh_tile;

/* Thread-local text of the last error. Returns the message length. */
int lsh_last_error(char* buffer, size_t capacity);

/* Library version string "major.minor.patch". */
const char* lsh_version(void);

/* Number of visible HIP devices (0 when none). Replaces the device probe of
 * cuda.py:14-138 / CudaSession.__init__ (cuda_backend.py:1160-1176). */
int lsh_device_count(int32_t* count);

/* Device name and compute-unit count. */
int lsh_device_info(int32_t device, char* name, size_t capacity, int32_t* compute_units);

/* Session: one device, one stream, resident DEMs and reusable scratch
 * (CudaSession, cuda_backend.py:1157-1258). */
int lsh_session_create(int32_t device, lsh_session** session);
int lsh_session_destroy(lsh_session* session);

/* Device memory helpers (the reference's cuda.device_array / to_device /
 * copy_to_host). Copies are synchronous with respect to the session stream. */
int lsh_alloc(lsh_session* session, size_t bytes, void** device_pointer);
int lsh_free(lsh_session* session, void* device_pointer);
int lsh_copy_to_device(lsh_session* session, void* device_dst, const void* host_src, size_t bytes);
int lsh_copy_to_host(lsh_session* session, void* host_dst, const void* device_src, size_t bytes);
int lsh_memset_device(lsh_session* session, void* device_dst, int value, size_t bytes);
int lsh_synchronize(lsh_session* session);

/* Timing events on the session stream (hipEvent): create, record at the
 * stream's current position, and read the milliseconds between two recorded
 * events (waits for the later one). No reference counterpart: the reference
 * times passes with cuda.event (cuda_backend.py:1491-1536). */
int lsh_event_create(lsh_session* session, void** event);
int lsh_event_record(lsh_session* session, void* event);
int lsh_event_elapsed_ms(void* start, void* stop, float* milliseconds);
int lsh_event_destroy(void* event);

/* Register a DEM on the device and build its factor-four max pyramid there
 * (CudaSession.build_max_pyramid, cuda_backend.py:1260-1281, and the resident
 * pyramids of _prepare_production_pyramids, cuda_backend.py:1231-1258).
 * elevation: host float32[height][width]; geo_transform: 6 float64
 * (origin_x, t1, t2, origin_y, t4, t5); projection: 6 float64 (radius,
 * lat0, lon0, scale, false_easting, false_northing). mips_host (nullable):
 * a cached pyramid payload (pyramid.py:34-48) used instead of building.
 * LSH_EINVAL when a side reaches 2^24 or when level 0, the mips and the
 * per-cell bound copy together exceed 4 GiB (the march addresses them with
 * 32-bit offsets). */
int lsh_dem_create(lsh_session* session, const float* elevation, int32_t width, int32_t height,
                   const double* geo_transform, const double* projection,
                   const float* mips_host, int32_t* dem_id);
int lsh_dem_destroy(lsh_session* session, int32_t dem_id);

/* Pyramid layout of a width x height DEM (pyramid.py:19-31): levels int32[L][4]
 * = (level, offset, width, height); capacity in rows. */
int lsh_pyramid_layout(int32_t width, int32_t height, int32_t* levels, int32_t capacity,
                       int32_t* level_count, int64_t* mip_count);

/* Copy a registered DEM's mips to the host (the .pyr.bin payload,
 * pyramid.py:56-75). mips: float32[mip_count]. */
int lsh_dem_pyramid(lsh_session* session, int32_t dem_id, float* mips_host, int64_t capacity);

/* Device pointers of a registered DEM (level 0 and mips), for callers that
 * build their own kernels' inputs. */
int lsh_dem_device_pointers(lsh_session* session, int32_t dem_id, const float** level0,
                            const float** mips);

/* FP64 ray-origin, direction, chord sampling and segment fitting of one tile on
 * the GPU (build_subpatch_segments_numba, geometry.py:575-673, with
 * generate_segments_parallel, geometry_numba.py:352-364).
 * segments (device): float32[N][C][D][18], C = (tile_width/subpatch + 2)^2. */
int lsh_segments_tile(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                      int32_t tile_column, int32_t tile_row, int32_t tile_width,
                      int32_t azimuth_count, double maximum_distance_m,
                      double observer_elevation_m, int32_t subpatch_size, float* segments);

/* Every DEM pass of one horizon tile in a single launch, the running maximum
 * carried in registers (CudaSession.subpatch_hierarchical_all_passes,
 * cuda_backend.py:1445-1552, and subpatch_hierarchy_kernel,
 * cuda_backend.py:1007-1134). segments (device): float32[N][C][D][18];
 * output (device): float32[tile_height*tile_width][N], pixel-major (or
 * [N][tile_height*tile_width] with LSH_HORIZON_AZIMUTH_MAJOR); pixels
 * outside the primary DEM are -inf. counters (device, nullable, with
 * LSH_HORIZON_COUNTERS): int64[4] += (level-0 samples, bound tests, steps,
 * 64 x the longest lane's steps per wavefront). */
int lsh_horizon_tile(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                     const float* segments, int32_t azimuth_count, int32_t tile_column,
                     int32_t tile_row, int32_t tile_width, int32_t tile_height,
                     int32_t subpatch_size, float observer_elevation_m, int32_t flags,
                     float* output, int64_t* counters);

/* Segments and march of one tile in one call (lsh_segments_tile followed by
 * lsh_horizon_tile) using session-owned segment scratch. */
int lsh_horizon_tile_full(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                          int32_t tile_column, int32_t tile_row, int32_t azimuth_count,
                          double maximum_distance_m, float observer_elevation_m,
                          int32_t flags, float* output, int64_t* counters);

/* Forward march of one 128x128 tile (as lsh_horizon_tile, subpatch 8, float64
 * heights) that also stores, per (pixel, azimuth), the int32 index of the
 * sample that set the horizon: pass << 24 | level-0 sample index within that
 * pass, or -1 when no sample exists. No reference counterpart: it feeds the
 * adjoint (the reference has no differentiation).
 * output float32 [16384][N] (device), winners int32 [16384][N] (device). */
int lsh_horizon_forward(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                        const float* segments, int32_t azimuth_count, int32_t tile_column,
                        int32_t tile_row, float observer_elevation_m, int32_t flags,
                        float* output, int32_t* winners);

/* Smoothed sun-fraction loss of horizon tiles and its gradient with respect to
 * the horizons (degrees): per pixel sum over Sun vectors of the 16-slice disk
 * fraction (lightmap.py:32-73) with each slice clamp rounded over
 * +-smoothing_deg (C1). loss (device, nullable) float64 [tile][16384];
 * grad_horizon (device, nullable) float32 [tile][16384][N]. Inputs as
 * lsh_lighting_tiles (pixel-major horizons), 2 <= N <= 1440; the slice walk
 * scales the reference's 1440-bin constants to N bins. smoothing_deg must not
 * exceed the narrowest slice half-height (0.094 deg for the reference disk). */
int lsh_sun_loss(lsh_session* session, int32_t dem_id, const lsh_tile* tiles,
                 int32_t tile_count, const float* horizons, int32_t azimuth_count,
                 const double* sun, int32_t sun_count, double sun_half_angle_deg,
                 double smoothing_deg, double* loss, float* grad_horizon);

/* Adjoint of lsh_horizon_forward for one tile: re-marches every ray whose
 * grad_horizon (float32 [16384][N], d loss / d horizon degrees) is non-zero to
 * its winning sample and adds d loss / d elevation to the four bilinear nodes of
 * that sample (in its DEM) and to the observer-terrain nodes of the pixel (in
 * the primary DEM). grad_fixed: host array of dem_count device int64 buffers
 * [height][width], accumulated as round(value * fixed_scale) with integer
 * atomics (deterministic). sample_cells (device, nullable) int32 [16384][N]:
 * pass << 28 | row << 14 | column of the winning sample's cell, -1 if skipped.
 * Rays are differentiated with the segment geometry held fixed. */
int lsh_horizon_adjoint(lsh_session* session, const int32_t* dem_ids, int32_t dem_count,
                        const float* segments, int32_t azimuth_count, int32_t tile_column,
                        int32_t tile_row, float observer_elevation_m, const int32_t* winners,
                        const float* grad_horizon, int64_t* const* grad_fixed,
                        double fixed_scale, int32_t* sample_cells);

/* Batched horizon generation: many tiles of many sites per launch.
 * It stands in for the reference's per-tile loop (generate_patch_horizons over
 * the bounded pipeline, generator.py:11-36, pipeline.py:127-450) when many
 * landing-site patches are generated together; the per-ray method and results
 * are those of lsh_horizon_tile, bit for bit.
 * A batch owns per-tile segment slots, the device site and tile tables, a
 * pinned staging copy of the tile list, the work-queue counters and a hipGraph
 * of the per-batch sequence (tile-list upload, FP64 segments of every tile, a
 * counter reset, the persistent march). */
typedef struct lsh_batch lsh_batch;

typedef struct {
  int32_t dem_ids[4];         /* the site's ordered DEM list, primary first */
  int32_t dem_count;          /* 1..4 */
  float observer_elevation_m; /* observer height above the primary DEM */
} lsh_batch_site;

typedef struct {
  int32_t site;        /* index into the batch's site list */
  int32_t tile_column; /* 128x128 tile origin in the site's primary DEM */
  int32_t tile_row;
  int32_t reserved;    /* must be 0 */
} lsh_batch_tile;

/* Kernel shapes the autotuner may choose for lsh_batch_run's `config`:
 * block_x x block_y pixels per wavefront (64 rays) and `azimuths` wavefronts
 * (consecutive azimuths) per workgroup. Shape 0 is the plain kernel's. */
int lsh_batch_shape_count(int32_t* count);
int lsh_batch_shape(int32_t config, int32_t* block_x, int32_t* block_y, int32_t* azimuths);

int lsh_batch_create(lsh_session* session, const lsh_batch_site* sites, int32_t site_count,
                     int32_t max_tiles, int32_t azimuth_count, double maximum_distance_m,
                     lsh_batch** batch);

/* The workgroup count at which every workgroup of a batch kernel shape is
 * resident at once on the session's device (the kernel's occupancy per CU x the
 * CU count, a multiple of 8): what lsh_batch_run uses when workgroups <= 0. */
int lsh_batch_resident_workgroups(lsh_session* session, int32_t config, int32_t* workgroups);

/* Horizons of tile_count <= max_tiles tiles into output (device float32
 * [tile_count][16384][azimuth_count], pixel-major per tile), in one persistent
 * launch: `workgroups` resident workgroups (a multiple of 8; <= 0 picks one
 * per the kernel's occupancy) drain per-XCD work queues of (azimuth group,
 * tile, pixel block) items, an XCD first taking the items of its own angular
 * sector and then stealing. flags: LSH_HORIZON_DEGREES and/or
 * LSH_HORIZON_AZIMUTH_MAJOR (each tile slot [N][16384]). use_graph = 1
 * replays the captured sequence (captured again when tile_count, config,
 * workgroups, flags or output change). Stream ordered; the call returns after
 * enqueueing (lsh_synchronize waits). */
int lsh_batch_run(lsh_batch* batch, const lsh_batch_tile* tiles, int32_t tile_count,
                  int32_t config, int32_t workgroups, int32_t flags, int32_t use_graph,
                  float* output);
int lsh_batch_destroy(lsh_batch* batch);

/* Lighting of stored horizon tiles in one streamed pass (config of the
 * reference's LightmapCudaSession.iter_patch_tiles / iter_patch_fraction_tiles /
 * iter_patch_margin_tiles, lightmap_cuda.py:187-276, and
 * PsrCudaSession.compute_patch, psr_cuda.py:238-375).
 * tiles (host): tile origins in the DEM; horizons (device): float32 degrees,
 * [tile][16384][N] (pixel-major, the tile file layout) or, with
 * azimuth_major = 1, [tile][N][16384]. sun / earth / psr_vectors (device):
 * float64 Moon-ME positions [count][3] (nullable when the count is 0).
 * sun_half_angle_deg: the disk half angle of the 16-slice sun (reference 0.27);
 * psr_diameter_deg: the PSR upper-limb diameter (reference 0.545). The slice
 * walk is the reference's at N = 1440 and scales its bin constants to N bins
 * otherwise (N >= 2).
 * Outputs (device, each nullable, selected by flags LSH_LIGHT_*):
 *   sun_bytes uint8 [tile][sun_count][16384] = uint8(255 * fraction)
 *   sun_fraction float [tile][sun_count][16384]; sun_mean float [tile][16384]
 *   earth_margin float [tile][earth_count][16384]  (elevation - horizon, deg)
 *   earth_visible float [tile][16384]  (fraction of times with margin >= 0)
 *   psr uint8 [tile][16384] (255 = PSR); psr_margin float [tile][16384]
 * Pixels outside the DEM get 0 / -inf. */
int lsh_lighting_tiles(lsh_session* session, int32_t dem_id, const lsh_tile* tiles,
                       int32_t tile_count, const float* horizons, int32_t azimuth_count,
                       int32_t azimuth_major, const double* sun, int32_t sun_count,
                       const double* earth, int32_t earth_count, const double* psr_vectors,
                       int32_t psr_count, double sun_half_angle_deg, double psr_diameter_deg,
                       int32_t flags, uint8_t* sun_bytes, float* sun_fraction, float* sun_mean,
                       float* earth_margin, float* earth_visible, uint8_t* psr,
                       float* psr_margin);

/* lsh_lighting_tiles without allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Arguments,
 * outputs and results are those of lsh_lighting_tiles; the device copy of the
 * tile list goes to a static table of the library instead of session scratch,
 * so the call performs no hipMalloc / hipFree. tile_count must be at most 4096
 * (LSH_EINVAL beyond). The table is one per device, shared by every call on
 * that device: calls on one stream are ordered, calls that overlap on one
 * device from different streams or sessions race. */
int lsh_lighting_tiles_static(lsh_session* session, int32_t dem_id, const lsh_tile* tiles,
                              int32_t tile_count, const float* horizons, int32_t azimuth_count,
                              int32_t azimuth_major, const double* sun, int32_t sun_count,
                              const double* earth, int32_t earth_count,
                              const double* psr_vectors, int32_t psr_count,
                              double sun_half_angle_deg, double psr_diameter_deg, int32_t flags,
                              uint8_t* sun_bytes, float* sun_fraction, float* sun_mean,
                              float* earth_margin, float* earth_visible, uint8_t* psr,
                              float* psr_margin);

/* The same lighting on host threads: the "cpu" backend of the lighting products
 * (LightmapCpuSession, lightmap_cpu.py:193-237; compute_psr_patch_reference,
 * psr.py:94-140). Every buffer is host memory: elevation float32
 * [height][width], geo_transform / projection float64[6] (as lsh_dem_create),
 * the other arguments and outputs as lsh_lighting_tiles. threads <= 0 uses
 * every core. */
int lsh_lighting_tiles_host(const float* elevation, int32_t width, int32_t height,
                            const double* geo_transform, const double* projection,
                            const lsh_tile* tiles, int32_t tile_count, const float* horizons,
                            int32_t azimuth_count, int32_t azimuth_major, const double* sun,
                            int32_t sun_count, const double* earth, int32_t earth_count,
                            const double* psr_vectors, int32_t psr_count,
                            double sun_half_angle_deg, double psr_diameter_deg, int32_t flags,
                            int32_t threads, uint8_t* sun_bytes, float* sun_fraction,
                            float* sun_mean, float* earth_margin, float* earth_visible,
                            uint8_t* psr, float* psr_margin);

/* Per-pixel reductions of the lighting time series (the downstream products'
 * stream reducers), computed in the same walk over the times as the signals,
 * without materialising the [time][pixel] stream:
 *   LSH_REDUCE_LONGEST_DURATION replaces reduce_longest_candidate_duration_stream
 *     (mission_duration.py:178-256) fed by the session's fraction / margin tiles
 *     as mission_duration_pipeline.py:267-318 feeds it: the condition
 *     signal >= sun_threshold [AND Earth margin >= earth_threshold_deg], float32
 *     comparisons; per candidate-start interval k (samples
 *     [interval_first[k], interval_end[k])) the longest duration in hours (days
 *     when output_days), summed from sample_hours in time order;
 *   LSH_REDUCE_SAFE_HAVEN replaces reduce_safe_haven_patch_stream
 *     (safe_haven.py:79-243): low Sun = fraction < sun_threshold, outage = Earth
 *     margin < earth_threshold_deg; per month band b (month_of[t] = b, or -1) the
 *     longest low-Sun run touching an outage in b, times step_hours, NaN where the
 *     band saw no outage or no Earth-above sample.
 * The sun signal is LSH_SIGNAL_SUN_FRACTION (the 16-slice disk fraction) or
 * LSH_SIGNAL_SUN_MARGIN (centre elevation - horizon, degrees); the safe haven
 * always uses the fraction. Earth margins need the earth vectors (same count). */
#define LSH_REDUCE_LONGEST_DURATION 0
#define LSH_REDUCE_SAFE_HAVEN 1
#define LSH_SIGNAL_SUN_FRACTION 0
#define LSH_SIGNAL_SUN_MARGIN 1

typedef struct {
  int32_t reducer;              /* LSH_REDUCE_* */
  int32_t sun_signal;           /* LSH_SIGNAL_* (longest duration) */
  double sun_threshold;         /* compared in float32 */
  int32_t use_earth;            /* longest duration: also require the Earth margin */
  double earth_threshold_deg;   /* compared in float32 */
  int32_t time_count;           /* T: sun (and earth) vector count */
  const double* sample_hours;   /* [T] hours each sample stands for (longest duration) */
  int32_t interval_count;       /* K candidate-start intervals (longest duration) */
  const int32_t* interval_first;  /* [K] first sample index inside interval k */
  const int32_t* interval_end;    /* [K] one past the last sample index inside k */
  int32_t output_days;          /* longest duration in days instead of hours */
  const int32_t* month_of;      /* [T] month band of each sample or -1 (safe haven) */
  int32_t band_count;           /* B month bands (safe haven) */
  double step_hours;            /* safe haven: hours per sample, applied in float32 */
} lsh_duration_spec;

/* Bytes of the workspace lsh_duration_tiles(_host) needs for tile_count tiles. */
int lsh_duration_workspace_bytes(const lsh_duration_spec* spec, int32_t tile_count,
                                 int64_t* bytes);

/* The reductions on the GPU (no reference C ABI: this replaces the numpy
 * reducers above and the per-time tiles that fed them). tiles (host), horizons
 * (device) and the vectors (device, float64 [T][3]; earth nullable when unused)
 * as lsh_lighting_tiles; the spec and its arrays are host memory (validated,
 * then copied to the device with the call); workspace (device) holds
 * lsh_duration_workspace_bytes and needs no initialisation; out (device)
 * float32 [band][tile][16384], bands = K or B; pixels outside the DEM get NaN.
 * The call returns once the reduction has finished. */
int lsh_duration_tiles(lsh_session* session, int32_t dem_id, const lsh_tile* tiles,
                       int32_t tile_count, const float* horizons, int32_t azimuth_count,
                       int32_t azimuth_major, const double* sun, const double* earth,
                       double sun_half_angle_deg, const lsh_duration_spec* spec,
                       void* workspace, float* out);

/* The same reductions on host threads (the "cpu" backend); every pointer,
 * including the spec's arrays and the workspace, is host memory. */
int lsh_duration_tiles_host(const float* elevation, int32_t width, int32_t height,
                            const double* geo_transform, const double* projection,
                            const lsh_tile* tiles, int32_t tile_count, const float* horizons,
                            int32_t azimuth_count, int32_t azimuth_major, const double* sun,
                            const double* earth, double sun_half_angle_deg,
                            const lsh_duration_spec* spec, int32_t threads, void* workspace,
                            float* out);

/* Five-viewpoint, per-azimuth-bin maximum-elevation reduction of Sun vectors
 * for PSR (reduce_sun_vectors_for_psr, psr.py:61-91). Host-only float64.
 * elevation: host float32[height][width]; vectors float64[count][3];
 * indices: int64[5 * azimuth_count] capacity, written in first-seen order. */
int lsh_psr_reduce(const float* elevation, int32_t width, int32_t height,
                   const double* geo_transform, const double* projection, const double* vectors,
                   int32_t count, int32_t azimuth_count, int64_t* indices, int32_t* kept);

/* Sun-disk slice table of the lightmap (lightmap.py:15-29) for a half angle:
 * half_circle float32[16] and max_photons. */
int lsh_sun_disk(double sun_half_angle_deg, float* half_circle, float* max_photons);

/* .cbin horizon tile codec (host only; no device is touched).
 * lsh_cbin_encode: the C# HorizonCompressor encoder (file_format.py:39-73,
 * _write_compressed :128-133): degrees float32 [pixels][azimuths] (host, no NaN)
 * -> payload (host, capacity bytes; pixels * (2 * azimuths + 2) always
 * suffices): per pixel a little-endian u16 block length and the block.
 * *payload_bytes receives the payload size. threads <= 0 uses every core.
 * lsh_cbin_decode: the decoder with its structural validation
 * (_decode_compressed_payload_into_python, file_format.py:135-182); LSH_EINVAL
 * names the first structural error. degrees: float32 [pixels][azimuths].
 * lsh_cbin_complete: HorizonTileStore.is_complete for a .cbin payload
 * (file_format.py:335-356): *complete = 1 when every block length is in
 * [1, 2 * azimuths], every block is present and nothing follows the last. */
int lsh_cbin_encode(const float* degrees, int32_t pixels, int32_t azimuths, uint8_t* payload,
                    int64_t capacity, int64_t* payload_bytes, int32_t threads);
int lsh_cbin_decode(const uint8_t* payload, int64_t bytes, int32_t pixels, int32_t azimuths,
                    float* degrees, int32_t threads);
int lsh_cbin_complete(const uint8_t* payload, int64_t bytes, int32_t pixels, int32_t azimuths,
                      int32_t* complete);

/* Horizon tile files (host only). The kind of a file is its extension (.bin
 * raw little-endian float32 [pixels][azimuths], .cbin compressed).
 * lsh_tile_file_status: HorizonTileStore.is_complete (file_format.py:335-356):
 * *status = LSH_TILE_ABSENT (no regular file), LSH_TILE_COMPLETE (a .bin of
 * exactly the raw size or raw + 28 legacy metadata bytes; a .cbin whose blocks
 * all parse with nothing after the last) or LSH_TILE_INCOMPLETE.
 * lsh_tile_file_read: read_horizon_tile_with_timings (file_format.py:219-274):
 * degrees float32 [pixels][azimuths]; seconds (nullable) double[2] = file read,
 * decode. LSH_EINVAL names an unreadable file, a wrong raw size or the first
 * structural error of a .cbin.
 * lsh_tile_file_write: HorizonTileStore.write's staged publish
 * (file_format.py:378-415): the raw values or the .cbin payload (compress = 1)
 * go to a new file beside `path`, are fsynced and renamed over `path`; on any
 * failure the staging file is removed and LSH_EIO returned. NaN is rejected
 * (LSH_EINVAL) before anything is written. The directory must exist. */
#define LSH_TILE_ABSENT 0
#define LSH_TILE_COMPLETE 1
#define LSH_TILE_INCOMPLETE 2
int lsh_tile_file_status(const char* path, int32_t pixels, int32_t azimuths, int32_t* status);
int lsh_tile_file_read(const char* path, int32_t pixels, int32_t azimuths, float* degrees,
                       int32_t threads, double* seconds);
int lsh_tile_file_write(const char* path, const float* degrees, int32_t pixels, int32_t azimuths,
                        int32_t compress, int32_t threads);

/* Measurement probes (no counterpart in the reference; they calibrate the
 * roofline of the march on the running device).
 * lsh_probe_copy: device-to-device copy of `bytes` (a multiple of 16) by a
 * streaming float4 kernel -- the known-size stream that scales rocprofv3's
 * FETCH_SIZE / WRITE_SIZE counters.
 * lsh_probe_gather: `threads` (a multiple of 256) threads each take `taps`
 * random 2x2 float32 taps of a width x height device window (width, height
 * >= 2) and write their maximum to out[threads]; the algorithmic gather rate is
 * 16 * threads * taps bytes over the kernel time. */
int lsh_probe_copy(lsh_session* session, void* device_dst, const void* device_src, size_t bytes);
int lsh_probe_gather(lsh_session* session, const float* window, int32_t width, int32_t height,
                     int32_t taps, int32_t threads, float* out);

#ifdef __cplusplus
}
#endif

#endif /* LUNARSCOUT_HIP_H */
