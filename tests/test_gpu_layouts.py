"""The azimuth-major horizon layout [azimuth][pixel] on the GPU.

The march writes it with coalesced stores (64 neighbouring pixels of one
azimuth per store instruction) and the lighting and duration kernels read it
with coalesced loads. Every path is checked bit for bit against the
pixel-major layout [pixel][azimuth] of the tile files:

* the plain march and the batched persistent march: the azimuth-major output is
  the transpose of the pixel-major one;
* lighting (every output: sun bytes and fractions, their mean, Earth margins and
  visibility, PSR mask and margin) and the fused duration reducers: identical
  outputs from either layout of the same horizons, on a full tile and a
  partial edge tile.
"""

from __future__ import annotations

import functools
from datetime import timedelta

import numpy as np
import pytest

from lunarscout_amd import durations as du
from lunarscout_amd import synthetic
from lunarscout_amd._lib import DeviceBuffer
from lunarscout_amd.geometry import DemGrid

pytestmark = pytest.mark.gpu

OUTPUTS = ("sun_bytes", "sun_fraction", "sun_mean", "earth_margin", "earth_visible", "psr",
           "psr_margin")


@functools.lru_cache(maxsize=None)
def session():
    from lunarscout_amd.hip_backend import HipSession

    return HipSession(0)


def _same(a, b):
    return (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b


def test_azimuth_major_march_is_the_transpose_bit_for_bit():
    s = session()
    dems = [synthetic.fbm_crater_dem(512, 20.0, 7)]
    for tile in ((128, 256), (384, 384)):
        segs = s.segments_device(dems, tile_column=tile[0], tile_row=tile[1])
        pm = s.march_device(dems, segs, tile_column=tile[0], tile_row=tile[1])
        am = s.march_device(dems, segs, tile_column=tile[0], tile_row=tile[1], azimuth_major=True)
        a, b = pm.to_host(), am.to_host()
        assert b.shape == (1440, 16384)
        same = (a.T == b) | (np.isneginf(a.T) & np.isneginf(b))
        print(f"tile {tile}: azimuth-major == transpose for {same.mean():.6f} of bins", flush=True)
        assert same.all()
        for buf in (segs, pm, am):
            buf.free()


def test_azimuth_major_batched_march_is_the_transpose_bit_for_bit():
    from lunarscout_amd.batch import HorizonBatch, Site

    s = session()
    regional = synthetic.fbm_crater_dem(1024, 60.0, 2)
    dem = synthetic.fbm_crater_dem(256, 10.0, 101, centre_xy=synthetic.site_xy())
    dem = DemGrid(np.ascontiguousarray(dem.elevation_m[:200, :256]), dem.geo_transform,
                  dem.projection)
    batch = HorizonBatch(s, [Site((dem, regional))], max_tiles=3)
    tiles = [(0, 0, 0), (0, 128, 0), (0, 0, 128)]  # the last one a partial edge tile
    pm = DeviceBuffer.empty(s, (3, 16384, 1440), np.float32)
    am = DeviceBuffer.empty(s, (3, 1440, 16384), np.float32)
    batch.run(tiles, pm, use_graph=False, degrees=True)
    batch.run(tiles, am, use_graph=False, degrees=True, azimuth_major=True)
    s.synchronize()
    a, b = pm.to_host(), am.to_host()
    batch.close()
    pm.free()
    am.free()
    for i in range(len(tiles)):
        same = (a[i].T == b[i]) | (np.isneginf(a[i].T) & np.isneginf(b[i]))
        print(f"batched tile {tiles[i]}: azimuth-major == transpose for {same.mean():.6f}",
              flush=True)
        assert same.all()


def _lit_case():
    dem = synthetic.fbm_crater_dem(384, 20.0, 4)  # tile (256, 256) is a partial edge tile
    tiles = [(128, 128), (256, 256)]
    s = session()
    horizons = np.stack([s.horizon_tile([dem], tile_column=c, tile_row=r, degrees=True)
                         for c, r in tiles])
    horizons = np.maximum(horizons, np.float32(-50.0))
    return dem, tiles, horizons


def test_lighting_reads_either_layout_to_the_same_bits():
    from lunarscout_amd.lighting import HipLightingSession, reduce_sun_vectors_for_psr

    dem, tiles, horizons = _lit_case()
    sun = synthetic.sun_vectors(dem, 500, 2.0, seed=8)
    earth = synthetic.earth_vectors(500, 2.0, seed=9)
    reduced, _ = reduce_sun_vectors_for_psr(dem, sun)
    lit = HipLightingSession(session())
    got = {}
    for major in (False, True):
        data = np.ascontiguousarray(horizons.transpose(0, 2, 1)) if major else horizons
        hz = DeviceBuffer.from_host(lit.session, data)
        out = lit.lighting_tiles(dem, tiles, hz, sun=sun, earth=earth, psr_vectors=reduced,
                                 azimuth_major=major, **{k: True for k in OUTPUTS})
        got[major] = {k: getattr(out, k).to_host() for k in OUTPUTS}
        out.free()
        hz.free()
    for k in OUTPUTS:
        same = _same(got[False][k], got[True][k])
        print(f"{k}: identical {same.mean():.6f} of {same.size}", flush=True)
        assert same.all(), k


def test_duration_reducers_read_either_layout_to_the_same_bits():
    from datetime import datetime, timezone

    from lunarscout_amd.lighting import HipLightingSession

    dem, tiles, horizons = _lit_case()
    T = 720
    sun = synthetic.sun_vectors(dem, T, 2.0, seed=10)
    earth = synthetic.earth_vectors(T, 2.0, seed=11)
    start = datetime(2027, 1, 1, tzinfo=timezone.utc)
    times = [start + timedelta(hours=2 * i) for i in range(T)]
    stop = times[-1] + timedelta(hours=2)
    plans = (
        du.longest_duration_plan(times, evaluation_start_utc=start, evaluation_stop_utc=stop,
                                 start_intervals=du.weekly_candidate_intervals(start, stop),
                                 sun_threshold=0.5, earth_threshold_deg=0.0),
        du.safe_haven_plan(times, sunlight_threshold=0.3, earth_threshold_deg=0.5,
                           time_step_hours=2.0),
    )
    lit = HipLightingSession(session())
    for plan in plans:
        got = []
        for major in (False, True):
            data = np.ascontiguousarray(horizons.transpose(0, 2, 1)) if major else horizons
            hz = DeviceBuffer.from_host(lit.session, data)
            out = lit.duration_tiles(dem, tiles, hz, plan, sun, earth, azimuth_major=major)
            got.append(out.to_host())
            out.free()
            hz.free()
        same = _same(got[0], got[1])
        print(f"reducer {plan.reducer}: identical {same.mean():.6f} of {same.size}", flush=True)
        assert same.all()
