"""Lightmap / PSR products over a stored horizon directory on the "cpu" backend.

The products are checked against the calculator they drive (CpuLightingSession,
itself checked against the C# captures and the float64 oracle in
test_lighting_cpu.py): every band and every patch lands where the reference
layout puts it, missing tiles carry invalid_value, and the public errors
follow the reference's codes.
"""

from __future__ import annotations

import numpy as np
import pytest

from lunarscout_amd import geotiff, synthetic, tiles
from lunarscout_amd.errors import InputError, OperationCancelledError, ProductGenerationError
from lunarscout_amd.lighting import CpuLightingSession, reduce_sun_vectors_for_psr
from lunarscout_amd.products import generate_lightmap, generate_psr


@pytest.fixture(scope="module")
def site(tmp_path_factory):
    root = tmp_path_factory.mktemp("site")
    dem = synthetic.fbm_crater_dem(256, 5.0, 1)
    dem = type(dem)(np.ascontiguousarray(dem.elevation_m[:200, :256]), dem.geo_transform, dem.projection)
    dem_path = geotiff.write_dem(root / "dem.tif", dem)
    store = tiles.HorizonTileStore(root / "horizons")
    rng = np.random.default_rng(2)
    az = np.radians(np.arange(1440) * 0.25)
    written = {}
    for ty, tx, vh, vw in ((0, 0, 128, 128), (0, 128, 128, 128), (128, 0, 72, 128)):
        hz = (rng.uniform(-1.5, 5.0, (vh * vw, 1)) + 1.2 * np.sin(2 * az)[None, :]).astype(np.float32)
        store.write(ty, tx, 0.0, hz, compress=True, valid_width=vw, valid_height=vh)
        written[(ty, tx)] = store.read(ty, tx, 0.0)
    # tile (128, 128) is missing on purpose
    sun = synthetic.sun_vectors(dem, 40, 6.0, seed=3)
    return dem, dem_path, root / "horizons", written, sun


def test_lightmap_bands_match_the_calculator(site, tmp_path):
    dem, dem_path, horizons, written, sun = site
    events = []
    out = generate_lightmap(dem_path, horizons, tmp_path / "lightmap.tif", sun_vectors_m=sun,
                            backend="cpu", invalid_value=7, progress_event_callback=events.append)
    assert events[0].stage == "start" and events[-1].stage == "complete"
    assert {e.backend for e in events} == {"cpu"}
    cpu = CpuLightingSession()
    for band in (0, 17, 39):
        values, info = geotiff.read_geotiff(out, band=band + 1)
        assert values.shape == (200, 256) and values.dtype == np.uint8
        for (ty, tx), hz in written.items():
            vh, vw = min(128, 200 - ty), min(128, 256 - tx)
            want = list(cpu.iter_patch_tiles(dem, hz, sun[band:band + 1], tile_y=ty, tile_x=tx,
                                             valid_height=vh, valid_width=vw))[0]
            np.testing.assert_array_equal(values[ty:ty + vh, tx:tx + vw], want)
        assert np.all(values[128:200, 128:256] == 7)


def test_psr_matches_the_calculator_and_missing_tiles_are_invalid(site, tmp_path):
    dem, dem_path, horizons, written, sun = site
    out = generate_psr(dem_path, horizons, tmp_path / "psr.tif", sun_vectors_m=sun, backend="cpu",
                       invalid_value=3)
    mask, _ = geotiff.read_geotiff(out)
    reduced, _ = reduce_sun_vectors_for_psr(dem, sun)
    cpu = CpuLightingSession()
    for (ty, tx), hz in written.items():
        vh, vw = min(128, 200 - ty), min(128, 256 - tx)
        np.testing.assert_array_equal(mask[ty:ty + vh, tx:tx + vw],
                                      cpu.compute_patch(dem, hz, reduced, tile_y=ty, tile_x=tx,
                                                        valid_height=vh, valid_width=vw))
    assert np.all(mask[128:200, 128:256] == 3)
    assert set(np.unique(mask[:128])) <= {0, 255}


def test_product_errors(site, tmp_path):
    dem, dem_path, horizons, written, sun = site
    cases = [
        (dict(horizons_path=tmp_path / "nope"), "product_horizons_not_found"),
        (dict(output_path=tmp_path / "x.png"), "product_output_extension_invalid"),
        (dict(backend="cuda"), "product_backend_invalid"),
        (dict(sun_vectors_m=None), "product_vectors_required"),
        (dict(sun_vectors_m=np.zeros((3, 2))), "product_vectors_invalid"),
        (dict(invalid_value=300), "product_invalid_value_invalid"),
    ]
    for change, code in cases:
        kwargs = dict(dem_path=dem_path, horizons_path=horizons, output_path=tmp_path / "p.tif",
                      sun_vectors_m=sun, backend="cpu")
        kwargs.update(change)
        with pytest.raises(InputError) as info:
            generate_psr(**kwargs)
        assert info.value.code == code, change
    (tmp_path / "exists.tif").write_bytes(b"x")
    with pytest.raises(ProductGenerationError) as info:
        generate_psr(dem_path, horizons, tmp_path / "exists.tif", sun_vectors_m=sun, backend="cpu")
    assert info.value.code == "product_output_exists"
    with pytest.raises(OperationCancelledError):
        generate_lightmap(dem_path, horizons, tmp_path / "c.tif", sun_vectors_m=sun, backend="cpu",
                          cancellation_requested=lambda: True)
    # cancelled: no output, the staged product stays for a resume (product_store.py:255-303)
    assert not (tmp_path / "c.tif").exists()
    assert len(list(tmp_path.glob(".c.tif.lunarscout-partial.*"))) == 3


def test_products_resume_after_cancellation(site, tmp_path):
    from lunarscout_amd.product_store import IncompatibleProductJobError

    dem, dem_path, horizons, written, sun = site
    calls = {"n": 0}

    def cancel_on_third_patch():
        calls["n"] += 1
        return calls["n"] > 20  # inside the second patch (one check per patch and per band)

    with pytest.raises(OperationCancelledError):
        generate_lightmap(dem_path, horizons, tmp_path / "r.tif", sun_vectors_m=sun[:12],
                          backend="cpu", cancellation_requested=cancel_on_third_patch)
    journal = tmp_path / ".r.tif.lunarscout-partial.journal.jsonl"
    done = journal.read_text().count("\n")
    assert 0 < done < 4
    # a torn last journal line is dropped on resume
    with open(journal, "a") as handle:
        handle.write('{"patch": "0,0", "sta')
    # another job's staged product is refused
    with pytest.raises(IncompatibleProductJobError) as info:
        generate_lightmap(dem_path, horizons, tmp_path / "r.tif", sun_vectors_m=sun[:11],
                          backend="cpu")
    assert info.value.code == "product_job_incompatible"
    events = []
    resumed = generate_lightmap(dem_path, horizons, tmp_path / "r.tif", sun_vectors_m=sun[:12],
                                backend="cpu", progress_event_callback=events.append)
    assert events[0].completed == done
    fresh = generate_lightmap(dem_path, horizons, tmp_path / "f.tif", sun_vectors_m=sun[:12],
                              backend="cpu")
    for band in (1, 7, 12):
        a, _ = geotiff.read_geotiff(resumed, band=band)
        b, _ = geotiff.read_geotiff(fresh, band=band)
        np.testing.assert_array_equal(a, b)
    assert not list(tmp_path.glob(".r.tif.lunarscout-partial.*"))
    # start_fresh discards a staged product of another job
    with pytest.raises(OperationCancelledError):
        generate_lightmap(dem_path, horizons, tmp_path / "s.tif", sun_vectors_m=sun[:5],
                          backend="cpu", cancellation_requested=lambda: True)
    out = generate_lightmap(dem_path, horizons, tmp_path / "s.tif", sun_vectors_m=sun[:12],
                            backend="cpu", start_fresh=True)
    a, _ = geotiff.read_geotiff(out, band=12)
    b, _ = geotiff.read_geotiff(fresh, band=12)
    np.testing.assert_array_equal(a, b)


def test_manifest_fields_tags_and_horizon_identity(site, tmp_path):
    """The job carries the reference's manifest fields and GeoTIFF tags
    (product_store.py:27-32, 154-252, 485-510, 624-640), and a horizon tile
    regenerated between a cancelled run and its resume changes the job."""
    import json
    import os
    import shutil
    from datetime import datetime, timedelta, timezone

    from lunarscout_amd.product_store import IncompatibleProductJobError

    dem, dem_path, horizons, written, sun = site
    store_dir = tmp_path / "horizons"
    shutil.copytree(horizons, store_dir)
    times = [datetime(2027, 3, 1, tzinfo=timezone.utc) + timedelta(hours=6 * i) for i in range(6)]
    calls = {"n": 0}

    def cancel_in_second_patch():
        calls["n"] += 1
        return calls["n"] > 10

    with pytest.raises(OperationCancelledError):
        generate_lightmap(dem_path, store_dir, tmp_path / "m.tif", sun_vectors_m=sun[:6],
                          times=times, backend="cpu", cancellation_requested=cancel_in_second_patch)
    manifest = json.loads((tmp_path / ".m.tif.lunarscout-partial.manifest.json").read_text())
    stamps = [t.strftime("%Y-%m-%dT%H:%M:%S.000000Z") for t in times]
    assert manifest["timestamps_utc"] == stamps
    assert manifest["band_count"] == 6 and len(manifest["band_metadata"]) == 6
    assert manifest["invalid_value"] == 0 and manifest["compute_backends"] == ["cpu"]
    assert manifest["horizon_inventory_identity"].startswith("stat-sha256:")
    # rewrite one horizon tile (same values, new file): the staged job no longer matches
    tile = next(store_dir.rglob("horizon_00000_00128_000.cbin"))
    payload = tile.read_bytes()
    tile.unlink()
    tile.write_bytes(payload)
    os.utime(tile, ns=(tile.stat().st_atime_ns, tile.stat().st_mtime_ns + 10**9))
    with pytest.raises(IncompatibleProductJobError):
        generate_lightmap(dem_path, store_dir, tmp_path / "m.tif", sun_vectors_m=sun[:6],
                          times=times, backend="cpu")
    out = generate_lightmap(dem_path, store_dir, tmp_path / "m.tif", sun_vectors_m=sun[:6],
                            times=times, backend="cpu", start_fresh=True)
    dataset, bands = geotiff.read_gdal_metadata(out)
    assert dataset["LUNARSCOUT_PRODUCT_SCHEMA"].startswith("lunarscout-")
    assert json.loads(dataset["LUNARSCOUT_TIMESTAMPS_UTC"]) == stamps
    assert json.loads(dataset["LUNARSCOUT_COMPUTE_BACKENDS"]) == ["cpu"]
    assert [bands[i]["TIMESTAMP_UTC"] for i in range(6)] == stamps
