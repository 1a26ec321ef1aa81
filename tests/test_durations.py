"""Mission-duration and safe-haven reductions: oracle pins, host logic, host build.

* the oracle reducers (oracle/reducers.py) against the known answers of the
  reference's own reducer tests (tests/numba_horizon/test_phase6b_mission_duration.py
  and test_phase6b_safe_haven.py: a candidate starting inside its interval,
  continuing past it and censored at the stop; irregular sample lengths with
  the inclusive threshold; the last sample credited to an unsampled stop; runs
  overlapping an outage credited in full; NaN for ill-posed months);
* the interval helpers against the reference's clipping / anchoring cases;
* the fused host reduction (lsh_duration_tiles_host) against the oracle
  reducers fed with the host lighting path's own per-time signal stacks: bit
  for bit (the reductions are exact given the signals);
* the products on the "cpu" backend, end to end through the GeoTIFF writer.
"""

from __future__ import annotations

from datetime import datetime, timedelta, timezone

import numpy as np
import pytest

from oracle import reducers
from lunarscout_amd import durations as du
from lunarscout_amd import synthetic
from lunarscout_amd.lighting import CpuLightingSession

UTC = timezone.utc


# -- oracle pins (the reference's known answers) ---------------------------------


def test_candidate_starts_inside_interval_continues_past_it_and_is_censored():
    times = [datetime(2027, 1, d, tzinfo=UTC) for d in range(1, 7)]
    intervals = [(times[1], times[3]), (times[3], times[5])]
    _, hours = du.validate_evaluation_samples(times, evaluation_start_utc=times[0],
                                              evaluation_stop_utc=times[-1])
    plan = du.longest_duration_plan(times, evaluation_start_utc=times[0],
                                    evaluation_stop_utc=times[-1], start_intervals=intervals,
                                    sun_threshold=0.5, output_unit="days")
    got = reducers.longest_candidate_durations(np.ones((6, 1, 1), bool), hours,
                                               plan.interval_first, plan.interval_end,
                                               output_days=True)
    assert got[:, 0, 0].tolist() == [4.0, 2.0]


def test_irregular_samples_and_inclusive_threshold():
    t0 = datetime(2027, 1, 1, tzinfo=UTC)
    times = [t0, t0 + timedelta(hours=6), t0 + timedelta(hours=15), t0 + timedelta(hours=24)]
    values = np.array([0.5, 0.5, 0.49, 1.0], np.float32)
    plan = du.longest_duration_plan(times, evaluation_start_utc=times[0],
                                    evaluation_stop_utc=times[-1],
                                    start_intervals=[(times[0], times[-1])], sun_threshold=0.5)
    got = reducers.longest_candidate_durations((values >= np.float32(0.5))[:, None, None],
                                               plan.sample_hours, plan.interval_first,
                                               plan.interval_end)
    assert got[0, 0, 0] == 15.0


def test_last_sample_is_credited_to_the_unsampled_stop():
    start = datetime(2027, 1, 1, tzinfo=UTC)
    stop = start + timedelta(hours=10)
    times = [start, start + timedelta(hours=6)]
    plan = du.longest_duration_plan(times, evaluation_start_utc=start, evaluation_stop_utc=stop,
                                    start_intervals=[(start, stop)], sun_threshold=0.5)
    assert plan.sample_hours.tolist() == [6.0, 4.0]
    got = reducers.longest_candidate_durations(np.ones((2, 1, 1), bool), plan.sample_hours,
                                               plan.interval_first, plan.interval_end)
    assert got[0, 0, 0] == 10.0


@pytest.mark.parametrize("fractions, earth, want", [
    ((0.1, 0.1, 0.1, 0.1), (1.0, 1.0, 1.0, 3.0), 24.0),   # run overlapping the outage, in full
    ((0.1, 0.1, 0.1, 0.8), (1.0, 3.0, 3.0, 3.0), 18.0),   # outage ends, run goes on
])
def test_safe_haven_credits_the_whole_overlapping_run(fractions, earth, want):
    times = [datetime(2027, 1, 1, h, tzinfo=UTC) for h in (0, 6, 12, 18)]
    plan = du.safe_haven_plan(times, sunlight_threshold=0.2, earth_threshold_deg=2.0,
                              time_step_hours=6.0)
    got = reducers.safe_haven(np.array(fractions, np.float32)[:, None, None],
                              np.array(earth, np.float32)[:, None, None], plan.month_of,
                              plan.band_count, sunlight_threshold=0.2, earth_threshold_deg=2.0,
                              time_step_hours=6.0)
    assert got.shape == (1, 1, 1) and got[0, 0, 0] == want


@pytest.mark.parametrize("earth", [(5.0, 5.0), (1.0, 1.0)])
def test_safe_haven_is_nan_when_ill_posed(earth):
    times = [datetime(2027, 1, 1, h, tzinfo=UTC) for h in (0, 6)]
    plan = du.safe_haven_plan(times, sunlight_threshold=0.2, earth_threshold_deg=2.0,
                              time_step_hours=6.0)
    got = reducers.safe_haven(np.full((2, 1, 1), 0.1, np.float32),
                              np.array(earth, np.float32)[:, None, None], plan.month_of, 1,
                              sunlight_threshold=0.2, earth_threshold_deg=2.0, time_step_hours=6.0)
    assert np.isnan(got[0, 0, 0])


# -- host interval logic --------------------------------------------------------


def test_interval_helpers_clip_months_and_anchor_weeks():
    months = du.monthly_candidate_intervals("2027-01-15T00:00:00Z", "2027-03-10T00:00:00Z")
    assert [(i.start_utc.day, i.stop_utc.day) for i in months] == [(15, 1), (1, 1), (1, 10)]
    weeks = du.weekly_candidate_intervals("2027-01-03T12:00:00Z", "2027-01-20T12:00:00Z")
    assert [i.stop_utc - i.start_utc for i in weeks] == [timedelta(days=7)] * 2 + [timedelta(days=3)]
    bands = du.build_month_bands(["2027-01-15T00:00:00Z", "2027-02-15T00:00:00Z"])
    assert [(a.month, b.month) for a, b in bands] == [(1, 2), (2, 3)]
    assert du.month_indices_map(["2027-01-01T00:00:00Z", "2027-02-01T00:00:00Z"],
                                bands).tolist() == [0, 1]


@pytest.mark.parametrize("kwargs, message", [
    (dict(start_intervals=[]), "at least one candidate-start interval"),
    (dict(start_intervals=[("2026-12-31T00:00:00Z", "2027-01-02T00:00:00Z")]), "inside the evaluation"),
    (dict(start_intervals=[("2027-01-01T00:00:00Z", "2027-01-02T00:00:00Z")], output_unit="weeks"),
     "output_unit"),
    (dict(start_intervals=[("2027-01-01T00:00:00Z", "2027-01-02T00:00:00Z")], sun_threshold=1.5),
     "between zero and one"),
])
def test_plan_validation_errors(kwargs, message):
    times = ["2027-01-01T00:00:00Z", "2027-01-01T12:00:00Z"]
    args = dict(evaluation_start_utc=times[0], evaluation_stop_utc="2027-01-03T00:00:00Z",
                sun_threshold=0.5)
    args.update(kwargs)
    with pytest.raises(ValueError, match=message):
        du.longest_duration_plan(times, **args)
    with pytest.raises(ValueError, match="UTC offset"):
        du.validate_evaluation_samples(["2027-01-01T00:00:00"], evaluation_start_utc=times[0],
                                       evaluation_stop_utc=times[1])


# -- the fused host reduction against the oracle on the host signal stacks ---------


def _site(T, step_hours=2.0):
    dem = synthetic.fbm_crater_dem(256, 20.0, 4)
    rng = np.random.default_rng(5)
    az = np.arange(1440, dtype=np.float32) * np.float32(0.25)
    base = rng.uniform(-1.0, 2.0, (16384, 1)).astype(np.float32)
    wave = rng.uniform(0.5, 2.5, (16384, 1)).astype(np.float32)
    phase = rng.uniform(0, 6.283, (16384, 1)).astype(np.float32)
    horizons = (base + wave * np.sin(np.radians(az)[None, :] * 3 + phase)).astype(np.float32)
    sun = synthetic.sun_vectors(dem, T, step_hours, seed=11, el_min=-2.0, el_max=3.5)
    earth = synthetic.earth_vectors(T, step_hours, seed=12)
    start = datetime(2027, 1, 1, tzinfo=UTC)
    times = [start + timedelta(hours=step_hours * i) for i in range(T)]
    return dem, horizons.reshape(1, 16384, 1440), sun, earth, times


def _signals(cpu, dem, horizons, sun, earth, tile):
    fr = cpu.lighting_tiles(dem, [tile], horizons, sun=sun, sun_mean=False, earth_visible=False,
                            psr=False, sun_fraction=True)["sun_fraction"][0]
    sun_margin = cpu.lighting_tiles(dem, [tile], horizons, earth=sun, sun_mean=False,
                                    earth_visible=False, psr=False,
                                    earth_margin=True)["earth_margin"][0]
    em = cpu.lighting_tiles(dem, [tile], horizons, earth=earth, sun_mean=False,
                            earth_visible=False, psr=False, earth_margin=True)["earth_margin"][0]
    return fr, sun_margin, em


def test_fused_host_longest_durations_equal_the_oracle_on_the_same_signals():
    T = 1080  # 90 days, 2-hour steps
    dem, horizons, sun, earth, times = _site(T)
    cpu = CpuLightingSession(threads=0)
    tile = (128, 0)
    fr, sm, em = _signals(cpu, dem, horizons, sun, earth, tile)
    stop = times[-1] + timedelta(hours=1)  # the last sample stands for one hour
    cases = [
        dict(start_intervals=du.weekly_candidate_intervals(times[0], stop), sun_signal="fraction",
             sun_threshold=0.5, earth_threshold_deg=None, output_unit="hours"),
        dict(start_intervals=[(times[10], times[400]), (times[200], times[900]),
                              (times[0], stop)],
             sun_signal="margin", sun_threshold=0.25, earth_threshold_deg=-1.0,
             output_unit="days"),
    ]
    for case in cases:
        plan = du.longest_duration_plan(times, evaluation_start_utc=times[0],
                                        evaluation_stop_utc=stop, **case)
        got = cpu.duration_tiles(dem, [tile], horizons, plan, sun, earth)[:, 0]
        signal = fr if case["sun_signal"] == "fraction" else sm
        cond = signal >= np.float32(case["sun_threshold"])
        if case["earth_threshold_deg"] is not None:
            cond &= em >= np.float32(case["earth_threshold_deg"])
        want = reducers.longest_candidate_durations(cond, plan.sample_hours, plan.interval_first,
                                                    plan.interval_end,
                                                    output_days=plan.output_days)
        assert 0.05 < cond.mean() < 0.95 and np.count_nonzero(want) > 1000
        np.testing.assert_array_equal(got, want)


def test_fused_host_safe_haven_equals_the_oracle_on_the_same_signals():
    T = 1200
    dem, horizons, sun, earth, times = _site(T)
    cpu = CpuLightingSession(threads=0)
    tile = (0, 128)
    fr, _, em = _signals(cpu, dem, horizons, sun, earth, tile)
    plan = du.safe_haven_plan(times, sunlight_threshold=0.3, earth_threshold_deg=0.5,
                              time_step_hours=2.0)
    assert plan.band_count == 4
    got = cpu.duration_tiles(dem, [tile], horizons, plan, sun, earth)[:, 0]
    want = reducers.safe_haven(fr, em, plan.month_of, plan.band_count, sunlight_threshold=0.3,
                               earth_threshold_deg=0.5, time_step_hours=2.0)
    finite = np.isfinite(want)
    assert 0.1 < finite.mean() and np.count_nonzero(want[finite]) > 1000
    np.testing.assert_array_equal(got, want)
    # a band table that is not in time order (the general path of the reduction)
    perm = np.array([2, 0, 3, 1], np.int32)
    shuffled = du.safe_haven_plan(times, sunlight_threshold=0.3, earth_threshold_deg=0.5,
                                  time_step_hours=2.0, month_bands=du.build_month_bands(times),
                                  month_index_of=perm[plan.month_of])
    got = cpu.duration_tiles(dem, [tile], horizons, shuffled, sun, earth)[:, 0]
    np.testing.assert_array_equal(got, want[np.argsort(perm)])


# -- products on the cpu backend -------------------------------------------------


def test_duration_products_write_the_reduced_bands(tmp_path):
    from lunarscout_amd.geotiff import read_geotiff
    from lunarscout_amd.tiles import HorizonTileStore

    T = 240
    dem, horizons, sun, earth, times = _site(T)
    store = HorizonTileStore(tmp_path / "h")
    for ty, tx in ((0, 0), (0, 128)):
        store.write(ty, tx, 0.0, horizons.reshape(16384, 1440), compress=False)
    stop = times[-1] + timedelta(hours=2)
    out = du.run_sunlight_earth_elevation_duration_product(
        dem=dem, horizon_store=store, output_path=tmp_path / "md.tif", times_utc=times,
        evaluation_start_utc=times[0], evaluation_stop_utc=stop,
        start_intervals=du.weekly_candidate_intervals(times[0], stop), sun_vectors_m=sun,
        earth_vectors_m=earth, sunlight_fraction_threshold=0.5,
        earth_elevation_threshold_deg=0.0, backend="cpu")
    plan = du.longest_duration_plan(times, evaluation_start_utc=times[0], evaluation_stop_utc=stop,
                                    start_intervals=du.weekly_candidate_intervals(times[0], stop),
                                    sun_threshold=0.5, earth_threshold_deg=0.0)
    cpu = CpuLightingSession()
    want = cpu.duration_patch(dem, horizons.reshape(128, 128, 1440), plan, sun, earth,
                              tile_y=0, tile_x=128)
    for band in range(plan.bands):
        data, _ = read_geotiff(out, band=band + 1)
        np.testing.assert_array_equal(data[0:128, 128:256], want[band])
        assert np.isnan(data[128:, :]).all()  # no horizon tile there: nodata
    sh = du.run_safe_haven_product(dem=dem, horizon_store=store, output_path=tmp_path / "sh.tif",
                                   times_utc=times, sun_vectors_m=sun, earth_vectors_m=earth,
                                   time_step_hours=2.0, backend="cpu")
    data, info = read_geotiff(sh, band=1)
    assert data.shape == (256, 256) and info.nodata is not None and np.isnan(info.nodata)
    with pytest.raises(Exception) as err:
        du.run_safe_haven_product(dem=dem, horizon_store=store, output_path=tmp_path / "sh.tif",
                                  times_utc=times, sun_vectors_m=sun, earth_vectors_m=earth,
                                  time_step_hours=2.0, backend="cpu")
    assert getattr(err.value, "code", None) == "product_output_exists"
    with pytest.raises(Exception) as err:
        du.run_safe_haven_product(dem=dem, horizon_store=store, output_path=tmp_path / "x.tif",
                                  times_utc=times, sun_vectors_m=sun[:-1], earth_vectors_m=earth,
                                  time_step_hours=2.0, backend="cpu")
    assert getattr(err.value, "code", None) == "product_vectors_invalid"


def test_body_elevation_products_write_one_margin_band_per_time(tmp_path):
    from lunarscout_amd.geotiff import read_geotiff
    from lunarscout_amd.tiles import HorizonTileStore

    T = 12
    dem, horizons, sun, earth, times = _site(T)
    store = HorizonTileStore(tmp_path / "h")
    store.write(0, 128, 0.0, horizons.reshape(16384, 1440), compress=False)
    out = du.run_earth_elevation_product(dem=dem, horizon_store=store,
                                         output_path=tmp_path / "earth.tif", times_utc=times,
                                         earth_vectors_m=earth, backend="cpu")
    cpu = CpuLightingSession()
    want = list(cpu.iter_patch_margin_tiles(dem, horizons.reshape(128, 128, 1440), earth,
                                            tile_y=0, tile_x=128))
    for t in (0, 5, T - 1):
        data, info = read_geotiff(out, band=t + 1)
        np.testing.assert_array_equal(data[0:128, 128:256], want[t])
        assert np.isnan(data[:, :128]).all()
    with pytest.raises(Exception) as err:
        du.run_sun_elevation_product(dem=dem, horizon_store=store, output_path=tmp_path / "s.tif",
                                     times_utc=times[:-1], sun_vectors_m=sun, backend="cpu")
    assert getattr(err.value, "code", None) == "product_vectors_invalid"
