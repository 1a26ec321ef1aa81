"""GPU parity of the horizon path: HIP kernels vs the float64 oracle.

Bars (north star): max |delta horizon| <= 0.01 deg and >= 99.9 % of bins within
1e-3 deg against the float64 oracle; pyramids bit-exact; the GPU march equal,
bit for bit, to the CPU build of the same kernel arithmetic on identical inputs.
The oracle runs on a seeded pixel subset of each tile (the GPU computes the
whole tile) so each case finishes in seconds.
"""

from __future__ import annotations

import functools

import numpy as np
import pytest

import oracle
from tests import fixtures as fx
from tests import kernel_cpu
from lunarscout_amd import synthetic
from lunarscout_amd.geometry import DemGrid, ProjectionParameters

pytestmark = pytest.mark.gpu

MAX_DEG = 0.01
BULK_DEG = 1e-3
BULK_FRACTION = 0.999


@functools.lru_cache(maxsize=None)
def session():
    from lunarscout_amd.hip_backend import HipSession

    return HipSession(0)


def degrees(slopes):
    return np.degrees(np.arctan(np.asarray(slopes, dtype=np.float64)))


def assert_horizon_parity(gpu_slopes, oracle_slopes, label):
    g, o = degrees(gpu_slopes), degrees(oracle_slopes)
    finite = np.isfinite(o)
    np.testing.assert_array_equal(np.isfinite(g), finite, err_msg=f"{label}: -inf pattern")
    err = np.abs(g - o)[finite]
    within = float(np.mean(err <= BULK_DEG)) if err.size else 1.0
    print(f"{label}: bins {err.size} max {err.max() if err.size else 0:.3g} deg, "
          f"within 1e-3 deg {within:.6f}", flush=True)
    assert err.size == 0 or err.max() <= MAX_DEG, label
    assert within >= BULK_FRACTION, label


def test_pyramid_is_bit_exact():
    rng = np.random.default_rng(11)
    for h, w in ((5, 7), (37, 53), (1024, 1024), (333, 1000)):
        z = rng.normal(0, 500, (h, w)).astype(np.float32)
        z[rng.random(z.shape) < 0.05] = np.nan
        z[rng.random(z.shape) < 0.02] = -25000.0
        z[rng.random(z.shape) < 0.01] = np.inf
        dem = DemGrid(z, np.array((0, 20, 0, 0, 0, -20), np.float64), ProjectionParameters.south_polar())
        got = session().build_max_pyramid(dem)
        levels, mips = oracle.build_pyramid(z)
        np.testing.assert_array_equal(got.levels, levels)
        assert got.mips.tobytes() == mips.tobytes(), (h, w)


def test_segments_match_captured_subpatch_tensor():
    meta = fx.rays_metadata()
    fixture = next(f for f in meta["subpatch_fixtures"] if f["id"] == "boundary_halo_multi_dem_16az")
    cfg = fixture["configuration"]
    dems = [fx.dem("multi_dem_outer_obstacle_east__dem_0"), fx.dem("multi_dem_outer_obstacle_east__dem_1")]
    got = session().build_subpatch_segments(
        dems, tile_column=cfg["tile_column"], tile_row=cfg["tile_row"], tile_width=cfg["tile_width"],
        azimuth_count=cfg["azimuth_count"], maximum_distance_m=cfg["max_distance_m"],
        observer_elevation_m=cfg["observer_elevation_m"], subpatch_size=cfg["subpatch_size"])
    expected = fx.rays_arrays()[f"{fixture['id']}__subpatch_fixture__segment_values"]
    np.testing.assert_allclose(got, expected, rtol=0, atol=1e-5)


def test_segments_match_oracle_on_polar_tile():
    dems = [synthetic.fbm_crater_dem(1024, 20.0, 0)]
    got = session().build_subpatch_segments(dems, tile_column=512, tile_row=512)
    ref = oracle.segments(dems, tile_column=512, tile_row=512, tile_width=128, azimuth_count=1440)
    diff = np.abs(got.astype(np.float64) - ref)
    tol = 1e-6 * np.maximum(np.abs(ref), 1.0)
    bad = diff > tol
    print("segments: fields beyond 1e-6 relative:", int(bad.sum()), "of", bad.size, flush=True)
    assert bad.mean() < 1e-5


def _tile_case(dems, tile, pixels=256, seed=0):
    rng = np.random.default_rng(seed)
    px = np.sort(rng.choice(16384, pixels, replace=False)).astype(np.int32)
    pyrs = [oracle.OraclePyramid(d) for d in dems]
    return pyrs, px


def test_march_equals_cpu_build_of_the_kernel_bit_for_bit():
    dems = [synthetic.fbm_crater_dem(512, 20.0, 3)]
    tile = (128, 256)
    segs = oracle.segments(dems, tile_column=tile[0], tile_row=tile[1], tile_width=128, azimuth_count=1440)
    pyrs, px = _tile_case(dems, tile, pixels=192)
    s = session()
    from lunarscout_amd._lib import DeviceBuffer

    seg_buf = DeviceBuffer.from_host(s, segs)
    out = s.march_device(dems, seg_buf, tile_column=tile[0], tile_row=tile[1]).to_host()
    cpu, _, _ = kernel_cpu.horizon_tile(pyrs, segs, tile_column=tile[0], tile_row=tile[1], pixels=px, mode=6)
    same = (out[px] == cpu) | (np.isneginf(out[px]) & np.isneginf(cpu))
    print("bit-identical fraction", same.mean(), flush=True)
    assert same.all()


def test_march_with_nodata_equals_cpu_build_and_oracle():
    """Two DEMs (1024^2 @5 m + 2048^2 @60 m) with nodata in both -- scattered NaN
    pixels, a -32768 block, an infinite pixel -- so rays take nodata skips,
    samples with invalid taps and guarded runs over nodata cells: the GPU equals
    the CPU build of the kernel bit for bit and the float64 oracle within the
    parity bar."""
    rng = np.random.default_rng(17)
    dems = []
    for d in synthetic.site_dems(synthetic.FULL_SITE, scale=2):
        z = np.array(d.elevation_m, dtype=np.float32)
        h, w = z.shape
        z[rng.random(z.shape) < 0.003] = np.nan
        z[h // 5:h // 5 + 120, w // 2:w // 2 + 140] = -32768.0
        z[h // 3, w // 4] = np.inf
        dems.append(DemGrid(z, d.geo_transform, d.projection))
    tile = (384, 384)
    segs = oracle.segments(dems, tile_column=tile[0], tile_row=tile[1], tile_width=128,
                           azimuth_count=1440)
    pyrs, px = _tile_case(dems, tile, pixels=192, seed=3)
    s = session()
    from lunarscout_amd._lib import DeviceBuffer

    seg_buf = DeviceBuffer.from_host(s, segs)
    out = s.march_device(dems, seg_buf, tile_column=tile[0], tile_row=tile[1]).to_host()
    cpu, _, _ = kernel_cpu.horizon_tile(pyrs, segs, tile_column=tile[0], tile_row=tile[1],
                                        pixels=px, mode=6)
    same = (out[px] == cpu) | (np.isneginf(out[px]) & np.isneginf(cpu))
    print("nodata site: bit-identical fraction", same.mean(), flush=True)
    assert same.all()
    ref, _ = oracle.horizon_tile(pyrs, segs, tile_column=tile[0], tile_row=tile[1], pixels=px,
                                 float64=True)
    assert_horizon_parity(out[px], ref, "nodata site tile")


def test_config1_tile_parity_against_float64_oracle():
    """1024x1024 px at 20 m, one DEM, central tile (512, 512), 1440 azimuths."""
    dems = [synthetic.fbm_crater_dem(1024, 20.0, 0)]
    tile = (512, 512)
    got = session().horizon_tile(dems, tile_column=tile[0], tile_row=tile[1])
    segs = oracle.segments(dems, tile_column=tile[0], tile_row=tile[1], tile_width=128, azimuth_count=1440)
    pyrs, px = _tile_case(dems, tile, pixels=256)
    ref, _ = oracle.horizon_tile(pyrs, segs, tile_column=tile[0], tile_row=tile[1], pixels=px, float64=True)
    assert_horizon_parity(got[px], ref, "config-1 tile")


def test_two_dem_tile_parity_against_float64_oracle():
    """Primary 5 m + regional 60 m (the full-site pair, grids shrunk 4x)."""
    dems = synthetic.site_dems(synthetic.FULL_SITE, scale=4)
    tile = (256, 128)
    got = session().horizon_tile(dems, tile_column=tile[0], tile_row=tile[1])
    segs = oracle.segments(dems, tile_column=tile[0], tile_row=tile[1], tile_width=128, azimuth_count=1440)
    pyrs, px = _tile_case(dems, tile, pixels=160, seed=1)
    ref, _ = oracle.horizon_tile(pyrs, segs, tile_column=tile[0], tile_row=tile[1], pixels=px, float64=True)
    assert_horizon_parity(got[px], ref, "two-DEM tile")


def test_partial_edge_tile_keeps_padding_at_minus_infinity():
    dems = [synthetic.fbm_crater_dem(1024, 20.0, 5)]
    crop = DemGrid(np.ascontiguousarray(dems[0].elevation_m[:300, :200]), dems[0].geo_transform,
                   dems[0].projection)
    tile = (128, 256)  # 72 columns x 44 rows inside the DEM
    got = session().horizon_tile([crop], tile_column=tile[0], tile_row=tile[1])
    grid = got.reshape(128, 128, 1440)
    assert np.isneginf(grid[44:]).all() and np.isneginf(grid[:, 72:]).all()
    segs = oracle.segments([crop], tile_column=tile[0], tile_row=tile[1], tile_width=128, azimuth_count=1440)
    rows, cols = np.meshgrid(np.arange(44), np.arange(72), indexing="ij")
    px = np.sort(np.random.default_rng(2).choice((rows * 128 + cols).ravel(), 96, replace=False))
    ref, _ = oracle.horizon_tile([oracle.OraclePyramid(crop)], segs, tile_column=tile[0], tile_row=tile[1],
                                 pixels=px.astype(np.int32), float64=True)
    assert_horizon_parity(got[px], ref, "edge tile")


def test_degrees_output_matches_slopes():
    dems = [synthetic.fbm_crater_dem(512, 20.0, 4)]
    s = session()
    slopes = s.horizon_tile(dems, tile_column=128, tile_row=128, azimuth_count=360)
    deg = s.horizon_tile(dems, tile_column=128, tile_row=128, azimuth_count=360, degrees=True)
    expected = np.degrees(np.arctan(slopes.astype(np.float64)))
    finite = np.isfinite(slopes)
    np.testing.assert_array_equal(np.isneginf(deg), ~finite)
    np.testing.assert_allclose(deg[finite], expected[finite], rtol=0, atol=2e-5)


def test_counters_match_the_cpu_build():
    dems = [synthetic.fbm_crater_dem(512, 20.0, 3)]
    segs = oracle.segments(dems, tile_column=128, tile_row=128, tile_width=128, azimuth_count=64)
    s = session()
    from lunarscout_amd._lib import DeviceBuffer, HORIZON_COUNTERS

    seg_buf = DeviceBuffer.from_host(s, segs)
    counters = DeviceBuffer.empty(s, (4,), np.int64)
    counters.zero()
    s.march_device(dems, seg_buf, tile_column=128, tile_row=128, flags=HORIZON_COUNTERS,
                   counters=counters)
    got = counters.to_host()
    _, _, expected = kernel_cpu.horizon_tile([oracle.OraclePyramid(d) for d in dems], segs,
                                             tile_column=128, tile_row=128, mode=6)
    np.testing.assert_array_equal(got[:3], expected)
    assert got[3] >= got[2]


def test_horizon_tile_returns_the_march_counters():
    dems = [synthetic.fbm_crater_dem(512, 20.0, 3)]
    s = session()
    _, got = s.horizon_tile(dems, tile_column=128, tile_row=128, azimuth_count=64, counters=True)
    segs = s.build_subpatch_segments(dems, tile_column=128, tile_row=128, azimuth_count=64)
    _, _, expected = kernel_cpu.horizon_tile([oracle.OraclePyramid(d) for d in dems], segs,
                                             tile_column=128, tile_row=128, mode=6)
    assert expected[0] > 0
    np.testing.assert_array_equal(got[:3], expected)
    assert got[3] >= got[2]


def test_captured_production_buffer_with_the_diagnostic_cap():
    """One pixel x 1440 azimuths x 2 passes of the captured C# buffer (cap on).

    The kernel's level-0 slope avoids the reference's float32 R + h rounding, so
    it is held to 2e-5 in slope; azimuth 360 (single-pixel spike) is excluded as
    in the oracle test.
    """
    base = "single_pixel_multi_dem_production__horizon_buffer_fixture"
    dems = [fx.dem(f"{base}__dem_{i}") for i in range(2)]
    s = session()
    segs = oracle.segments(dems, tile_column=20, tile_row=20, tile_width=1, azimuth_count=1440,
                           exact_chain=True)
    from lunarscout_amd._lib import DeviceBuffer, HORIZON_CAP

    seg_buf = DeviceBuffer.from_host(s, segs)
    out = s.march_device(dems, seg_buf, tile_column=20, tile_row=20, tile_width=1, tile_height=1,
                         flags=HORIZON_CAP).to_host()[0]
    expected = fx.rays_arrays()[f"{base}__final_slopes"][0]
    others = np.arange(1440) != 360
    np.testing.assert_allclose(out[others], expected[others], rtol=0, atol=2e-5)


def test_session_seam_takes_reference_pyramid_arrays_as_the_generator_calls_it():
    """generate_patch_horizons (generator.py:27-36) with the reference's PyramidArrays.

    The seam gets the captured subpatch segment tensor and pyramid look-alikes
    built from the fixture arrays PyramidArrays.from_artifact reads; its slopes
    must equal the plain march of the same segments over the fixture DEMs bit
    for bit, and the oracle within the parity bar.
    """
    from lunarscout_amd._lib import DeviceBuffer
    from tests.test_seams import reference_pyramid

    meta = fx.rays_metadata()
    fixture = next(f for f in meta["subpatch_fixtures"] if f["id"] == "boundary_halo_multi_dem_16az")
    cfg = fixture["configuration"]
    prefixes = ["multi_dem_outer_obstacle_east__dem_0", "multi_dem_outer_obstacle_east__dem_1"]
    values = np.ascontiguousarray(
        fx.rays_arrays()[f"{fixture['id']}__subpatch_fixture__segment_values"], dtype=np.float32)
    s = session()
    slopes, seconds = s.subpatch_hierarchical_all_passes(
        values, [reference_pyramid(p) for p in prefixes], tile_column=cfg["tile_column"],
        tile_row=cfg["tile_row"], tile_width=cfg["tile_width"], tile_height=cfg["tile_height"],
        subpatch_size=cfg["subpatch_size"], observer_elevation_m=cfg["observer_elevation_m"])
    assert slopes.shape == (cfg["tile_width"] * cfg["tile_height"], cfg["azimuth_count"])
    assert len(seconds) == 1
    dems = [fx.dem(p) for p in prefixes]
    seg = DeviceBuffer.from_host(s, values)
    plain = s.march_device(dems, seg, tile_column=cfg["tile_column"], tile_row=cfg["tile_row"],
                           tile_width=cfg["tile_width"], tile_height=cfg["tile_height"],
                           subpatch_size=cfg["subpatch_size"],
                           observer_elevation_m=cfg["observer_elevation_m"]).to_host()
    seg.free()
    same = (plain == slopes) | (np.isneginf(plain) & np.isneginf(slopes))
    assert same.all()
    ref, _ = oracle.horizon_tile([oracle.OraclePyramid(d) for d in dems], values,
                                 tile_column=cfg["tile_column"], tile_row=cfg["tile_row"],
                                 tile_width=cfg["tile_width"], tile_height=cfg["tile_height"],
                                 subpatch_size=cfg["subpatch_size"],
                                 observer_elevation_m=cfg["observer_elevation_m"], float64=True)
    assert_horizon_parity(slopes, ref, "seam with reference PyramidArrays")
