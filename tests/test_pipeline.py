"""Tile pipeline, resume rule, cancellation and the public generator's validation (CPU).

Mirrors the reference's pipeline tests (tests/numba_horizon/test_phase6_pipeline.py,
tests/test_public_horizon.py): a fake processor stands in for the GPU session,
the way the reference's FakeSession does (test_phase4_subpatch.py:140-157).
"""

from __future__ import annotations

import threading

import numpy as np
import pytest

from lunarscout_amd import geotiff, synthetic, tiles
from lunarscout_amd.errors import HipError, InputError, OperationCancelledError
from lunarscout_amd.horizon import generate_horizons
from lunarscout_amd.pipeline import (HorizonPipelineCancelled, enumerate_patches,
                                     run_bounded_pipeline)


def test_enumerate_patches_keeps_partial_edges():
    patches = enumerate_patches(300, 130)
    assert [(p.tile_x, p.tile_y, p.width, p.height) for p in patches] == [
        (0, 0, 128, 128), (128, 0, 128, 128), (256, 0, 44, 128),
        (0, 128, 128, 2), (128, 128, 128, 2), (256, 128, 44, 2)]
    assert [p.index for p in patches] == list(range(6))
    with pytest.raises(ValueError):
        enumerate_patches(300, 130, include_partial_edges=False)
    with pytest.raises(ValueError):
        enumerate_patches(0, 10)


def _fake_factory(calls, fail_on=None):
    def factory(worker_id):
        def process(patch, prepared):
            if fail_on is not None and (patch.tile_x, patch.tile_y) == fail_on:
                raise RuntimeError("simulated device failure")
            calls.append((patch.tile_x, patch.tile_y))
            value = np.float32(patch.tile_x + patch.tile_y / 1000.0) / np.float32(100.0)
            return np.full((16384, 1440), value, np.float32)
        return process
    return factory


@pytest.mark.parametrize("writer_queue", [None, 1])
def test_pipeline_writes_every_tile_and_resumes(tmp_path, writer_queue):
    store = tiles.HorizonTileStore(tmp_path)
    patches = enumerate_patches(300, 130)
    calls, events = [], []
    result = run_bounded_pipeline(patches, store=store, prepare_patch=lambda p: p.index,
                                  processor_factory=_fake_factory(calls), compress=True,
                                  writer_queue_capacity=writer_queue, progress_callback=events.append)
    assert len(result.output_paths) == 6 and result.skipped_patches == 0
    assert sorted(calls) == sorted((p.tile_x, p.tile_y) for p in patches)
    assert [e.stage for e in events][:2] == ["prepare_patches", "process_patches"]
    assert events[-1].stage == "complete" and events[-1].percent == 100.0
    edge = store.read(128, 256, 0.0)
    assert np.all(edge[2:] == np.float32(-50.0)) and np.all(edge[:, 44:] == np.float32(-50.0))
    # resume: complete tiles are skipped; a truncated one is regenerated
    victim = store.find_existing_path(0, 128, 0.0)
    victim.write_bytes(victim.read_bytes()[:-3])
    calls.clear()
    result = run_bounded_pipeline(patches, store=store, prepare_patch=lambda p: None,
                                  processor_factory=_fake_factory(calls), compress=True)
    assert calls == [(128, 0)] and result.skipped_patches == 5
    calls.clear()
    run_bounded_pipeline(patches, store=store, prepare_patch=lambda p: None,
                         processor_factory=_fake_factory(calls), compress=True, skip_existing=False)
    assert len(calls) == 6


def test_pipeline_propagates_the_first_error_and_keeps_complete_tiles(tmp_path):
    store = tiles.HorizonTileStore(tmp_path)
    patches = enumerate_patches(512, 256)
    calls = []
    with pytest.raises(RuntimeError, match="simulated"):
        run_bounded_pipeline(patches, store=store, prepare_patch=lambda p: None,
                             processor_factory=_fake_factory(calls, fail_on=(256, 0)),
                             compress=False, writer_queue_capacity=1)
    for p in patches:
        path = store.find_existing_path(p.tile_y, p.tile_x, 0.0, require_complete=False)
        assert path is None or store.is_complete(path)
    assert not list(tmp_path.rglob("*.tmp*"))


def test_pipeline_cancellation_between_tiles(tmp_path):
    store = tiles.HorizonTileStore(tmp_path)
    patches = enumerate_patches(512, 512)
    calls = []
    flag = threading.Event()

    def factory(worker_id):
        inner = _fake_factory(calls)(worker_id)

        def process(patch, prepared):
            if len(calls) == 3:
                flag.set()
            return inner(patch, prepared)
        return process

    with pytest.raises(HorizonPipelineCancelled):
        run_bounded_pipeline(patches, store=store, prepare_patch=lambda p: None,
                             processor_factory=factory, compress=True, writer_queue_capacity=1,
                             cancellation_requested=flag.is_set)
    assert 3 <= len(calls) < len(patches)
    written = [p for p in patches if store.find_existing_path(p.tile_y, p.tile_x, 0.0)]
    assert len(written) < len(patches)


def test_generate_horizons_validation(tmp_path):
    dem_path = geotiff.write_dem(tmp_path / "dem.tif", synthetic.fbm_crater_dem(256, 20.0, 1))
    cases = [
        (dict(dem_paths=str(dem_path)), "horizon_dem_paths_invalid"),
        (dict(dem_paths=[]), "horizon_dem_paths_empty"),
        (dict(dem_paths=[tmp_path / "missing.tif"]), "horizon_dem_not_found"),
        (dict(dem_paths=[dem_path], observer_height_m=100.0), "horizon_observer_height_invalid"),
        (dict(dem_paths=[dem_path], observer_height_m=float("nan")), "horizon_observer_height_invalid"),
        (dict(dem_paths=[dem_path], compress=1), "horizon_option_invalid"),
        (dict(dem_paths=[dem_path], progress_callback=3), "horizon_callback_invalid"),
    ]
    for kwargs, code in cases:
        with pytest.raises(InputError) as info:
            generate_horizons(tmp_path / "out", **kwargs)
        assert info.value.code == code, kwargs
    (tmp_path / "file").write_text("x")
    with pytest.raises(InputError) as info:
        generate_horizons(tmp_path / "file", [dem_path])
    assert info.value.code == "horizon_output_is_file"
    with pytest.raises(OperationCancelledError) as info:
        generate_horizons(tmp_path / "out", [dem_path], cancellation_requested=lambda: True)
    assert info.value.code == "horizon_generation_cancelled"
    assert not (tmp_path / "out").exists()


def test_generate_horizons_without_a_device_touches_no_tile(tmp_path):
    from lunarscout_amd import _lib

    if _lib.device_count() > 0:
        pytest.skip("a HIP device is present; the GPU tests cover generation")
    dem_path = geotiff.write_dem(tmp_path / "dem.tif", synthetic.fbm_crater_dem(256, 20.0, 1))
    with pytest.raises(HipError) as info:
        generate_horizons(tmp_path / "out", [dem_path])
    assert info.value.code == "hip_horizon_unavailable"
    assert not (tmp_path / "out").exists()


def test_geotiff_round_trip_of_dems_and_products(tmp_path):
    dem = synthetic.fbm_crater_dem(300, 20.0, 3)
    back = geotiff.load_dem(geotiff.write_dem(tmp_path / "a.tif", dem))
    np.testing.assert_array_equal(back.elevation_m, dem.elevation_m)
    np.testing.assert_array_equal(back.geo_transform, dem.geo_transform)
    assert back.projection == dem.projection
    for dtype in (np.uint8, np.int16, np.float64):
        v = (np.arange(130 * 131) % 251).astype(dtype).reshape(130, 131)
        path = geotiff.write_geotiff(tmp_path / f"p_{np.dtype(dtype).name}.tif", v,
                                     dem.geo_transform, dem.projection, nodata=0)
        values, info = geotiff.read_geotiff(path)
        np.testing.assert_array_equal(values, v)
        assert values.dtype == np.dtype(dtype) and info.nodata == 0.0
    with pytest.raises(InputError):
        geotiff.write_geotiff(tmp_path / "p_uint8.tif", v, dem.geo_transform, dem.projection)
