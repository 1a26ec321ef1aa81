"""The fused mission-duration and safe-haven reductions on the GPU.

* bit for bit against the oracle reducers (oracle/reducers.py) fed with the
  GPU's own per-time signal stacks (the lighting kernel's sun fraction / margin
  outputs): the reduction itself is exact;
* against the host build (the "cpu" backend): identical except where a signal
  lies within float32 rounding of its threshold (host and device math
  libraries differ in the last bit of atan2 / asin), at most 0.1 % of values;
* the hip product equals the cpu product under the same bar.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import reducers
from lunarscout_amd import durations as du
from lunarscout_amd._lib import DeviceBuffer
from lunarscout_amd.lighting import CpuLightingSession, HipLightingSession
from tests.test_durations import _site

pytestmark = pytest.mark.gpu


def _gpu_signals(light, dem, hz, sun, earth, tile):
    fr = light.lighting_tiles(dem, [tile], hz, sun=sun, sun_mean=False, earth_visible=False,
                              psr=False, sun_fraction=True).sun_fraction.to_host()[0]
    sm = light.lighting_tiles(dem, [tile], hz, earth=sun, sun_mean=False, earth_visible=False,
                              psr=False, earth_margin=True).earth_margin.to_host()[0]
    em = light.lighting_tiles(dem, [tile], hz, earth=earth, sun_mean=False, earth_visible=False,
                              psr=False, earth_margin=True).earth_margin.to_host()[0]
    return fr, sm, em


def test_gpu_reductions_equal_the_oracle_on_the_gpu_signals():
    from datetime import timedelta

    T = 2190  # half a year, 2-hour steps
    dem, horizons, sun, earth, times = _site(T)
    light = HipLightingSession()
    hz = DeviceBuffer.from_host(light.session, horizons)
    tile = (128, 128)
    fr, sm, em = _gpu_signals(light, dem, hz, sun, earth, tile)
    stop = times[-1] + timedelta(hours=2)
    plan = du.longest_duration_plan(times, evaluation_start_utc=times[0], evaluation_stop_utc=stop,
                                    start_intervals=du.monthly_candidate_intervals(times[0], stop),
                                    sun_signal="fraction", sun_threshold=0.5,
                                    earth_threshold_deg=0.0)
    got = light.duration_tiles(dem, [tile], hz, plan, sun, earth).to_host()[:, 0]
    cond = (fr >= np.float32(0.5)) & (em >= np.float32(0.0))
    want = reducers.longest_candidate_durations(cond, plan.sample_hours, plan.interval_first,
                                                plan.interval_end)
    print(f"longest durations: {plan.bands} bands, {np.count_nonzero(want)} non-zero", flush=True)
    np.testing.assert_array_equal(got, want)

    plan = du.longest_duration_plan(times, evaluation_start_utc=times[0], evaluation_stop_utc=stop,
                                    start_intervals=du.weekly_candidate_intervals(times[0], stop),
                                    sun_signal="margin", sun_threshold=0.3, output_unit="days")
    got = light.duration_tiles(dem, [tile], hz, plan, sun).to_host()[:, 0]
    want = reducers.longest_candidate_durations(sm >= np.float32(0.3), plan.sample_hours,
                                                plan.interval_first, plan.interval_end,
                                                output_days=True)
    np.testing.assert_array_equal(got, want)

    plan = du.safe_haven_plan(times, sunlight_threshold=0.3, earth_threshold_deg=0.5,
                              time_step_hours=2.0)
    got = light.duration_tiles(dem, [tile], hz, plan, sun, earth).to_host()[:, 0]
    want = reducers.safe_haven(fr, em, plan.month_of, plan.band_count, sunlight_threshold=0.3,
                               earth_threshold_deg=0.5, time_step_hours=2.0)
    print(f"safe haven: {plan.bands} bands, {np.isfinite(want).mean():.3f} posed", flush=True)
    np.testing.assert_array_equal(got, want)
    hz.free()


def _agreement(a, b):
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    return same.mean()


def test_gpu_and_host_builds_agree():
    from datetime import timedelta

    T = 1440
    dem, horizons, sun, earth, times = _site(T)
    light, cpu = HipLightingSession(), CpuLightingSession()
    hz = DeviceBuffer.from_host(light.session, horizons)
    tiles = [(0, 0)]
    stop = times[-1] + timedelta(hours=2)
    for plan in (
        du.longest_duration_plan(times, evaluation_start_utc=times[0], evaluation_stop_utc=stop,
                                 start_intervals=du.weekly_candidate_intervals(times[0], stop),
                                 sun_threshold=0.5, earth_threshold_deg=0.0),
        du.safe_haven_plan(times, sunlight_threshold=0.3, earth_threshold_deg=0.5,
                           time_step_hours=2.0),
    ):
        g = light.duration_tiles(dem, tiles, hz, plan, sun, earth).to_host()
        c = cpu.duration_tiles(dem, tiles, horizons, plan, sun, earth)
        agree = _agreement(g, c)
        print(f"reducer {plan.reducer}: gpu/host agreement {agree:.6f}", flush=True)
        assert agree >= 0.999
    hz.free()


def test_hip_product_matches_cpu_product(tmp_path):
    from datetime import timedelta

    from lunarscout_amd.geotiff import read_geotiff
    from lunarscout_amd.tiles import HorizonTileStore

    T = 360
    dem, horizons, sun, earth, times = _site(T)
    store = HorizonTileStore(tmp_path / "h")
    store.write(0, 0, 0.0, horizons.reshape(16384, 1440), compress=False)
    stop = times[-1] + timedelta(hours=2)
    common = dict(dem=dem, horizon_store=store, times_utc=times, evaluation_start_utc=times[0],
                  evaluation_stop_utc=stop,
                  start_intervals=du.weekly_candidate_intervals(times[0], stop),
                  sun_vectors_m=sun, sunlight_fraction_threshold=0.5)
    a = du.run_sunlight_duration_product(output_path=tmp_path / "hip.tif", backend="hip", **common)
    b = du.run_sunlight_duration_product(output_path=tmp_path / "cpu.tif", backend="cpu", **common)
    for band in (1, 2):
        x, _ = read_geotiff(a, band=band)
        y, _ = read_geotiff(b, band=band)
        assert _agreement(x, y) >= 0.999
