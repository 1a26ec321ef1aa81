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


def _bands(path, count):
    return [geotiff.read_geotiff(path, band=b)[0] for b in range(1, count + 1)]


def test_bounded_pipeline_writes_what_the_serial_one_writes(site, tmp_path):
    """Overlapped reads, calculation and batched durable writes change nothing
    in the product (psr_pipeline.py:523-797 against its serial mode)."""
    dem, dem_path, horizons, written, sun = site
    runs = {
        "serial": dict(pipeline_mode="serial", durable_batch_size=1),
        "bounded": dict(pipeline_mode="bounded", reader_worker_count=3,
                        decoded_horizon_capacity=2, writer_queue_capacity=1, durable_batch_size=2),
        "one-reader": dict(pipeline_mode="bounded", reader_worker_count=1,
                           decoded_horizon_capacity=1, durable_batch_size=16),
    }
    for name, product, count in (("lightmap", generate_lightmap, 12), ("psr", generate_psr, 1)):
        got = {}
        for label, knobs in runs.items():
            events = []
            out = product(dem_path, horizons, tmp_path / f"{name}-{label}.tif",
                          sun_vectors_m=sun[:12], backend="cpu", invalid_value=9,
                          progress_event_callback=events.append, **knobs)
            got[label] = _bands(out, count)
            states = [e.stage for e in events]
            assert states[0] == "start" and states[-1] == "complete"
            assert states.count("valid") == 3 and states.count("invalid") == 1, label
            durable = [e.completed for e in events if e.stage in ("valid", "invalid")]
            assert durable == [1, 2, 3, 4], label
        for label in ("bounded", "one-reader"):
            for a, b in zip(got["serial"], got[label]):
                np.testing.assert_array_equal(a, b, err_msg=f"{name} {label}")


def test_bounded_pipeline_resumes_after_cancellation_from_whole_checkpoints(site, tmp_path):
    import json

    dem, dem_path, horizons, written, sun = site
    calls = {"n": 0}

    def cancel_late():
        calls["n"] += 1
        return calls["n"] > 30  # inside the third patch's bands

    with pytest.raises(OperationCancelledError):
        generate_lightmap(dem_path, horizons, tmp_path / "b.tif", sun_vectors_m=sun[:12],
                          backend="cpu", cancellation_requested=cancel_late,
                          durable_batch_size=2, reader_worker_count=2)
    journal = tmp_path / ".b.tif.lunarscout-partial.journal.jsonl"
    entries = [json.loads(line) for line in journal.read_text().splitlines()]
    # every journaled patch is whole: all 12 bands, each block inside the synced end
    assert 1 <= len(entries) < 4
    for e in entries:
        assert len(e["blocks"]) == 12
        assert all(offset + count <= e["end"] for _, offset, count in e["blocks"])
    resumed = generate_lightmap(dem_path, horizons, tmp_path / "b.tif", sun_vectors_m=sun[:12],
                                backend="cpu", durable_batch_size=2)
    fresh = generate_lightmap(dem_path, horizons, tmp_path / "bf.tif", sun_vectors_m=sun[:12],
                              backend="cpu", pipeline_mode="serial")
    for a, b in zip(_bands(resumed, 12), _bands(fresh, 12)):
        np.testing.assert_array_equal(a, b)


def test_output_transform_converts_valid_patches_only(site, tmp_path):
    """products.py:361-372 and product_store.py:362-385: the conversion maps
    valid band blocks to output_dtype; invalid patches keep invalid_value; the
    transform's identity is part of the job."""
    from lunarscout_amd.product_store import IncompatibleProductJobError

    dem, dem_path, horizons, written, sun = site
    plain = generate_psr(dem_path, horizons, tmp_path / "plain.tif", sun_vectors_m=sun,
                         backend="cpu", invalid_value=3)
    scaled = generate_psr(dem_path, horizons, tmp_path / "scaled.tif", sun_vectors_m=sun,
                          backend="cpu", invalid_value=3,
                          output_transform=lambda b: (b.astype(np.float32) / 255.0),
                          output_dtype=np.float32, output_transform_id="unit-interval")
    a, _ = geotiff.read_geotiff(plain)
    b, _ = geotiff.read_geotiff(scaled)
    assert b.dtype == np.float32
    valid = np.ones(a.shape, bool)
    valid[128:200, 128:256] = False
    np.testing.assert_array_equal(b[valid], a[valid].astype(np.float32) / 255.0)
    assert np.all(b[~valid] == 3.0)
    bad = [
        (dict(output_dtype=np.float32), "product_output_conversion_invalid"),
        (dict(output_transform_id="x"), "product_output_conversion_invalid"),
        (dict(output_transform=lambda b: b), "product_output_conversion_invalid"),
        (dict(output_transform=3, output_dtype=np.uint8), "product_output_conversion_invalid"),
        (dict(output_transform=lambda b: b, output_dtype="not-a-dtype"),
         "product_output_dtype_invalid"),
        (dict(pipeline_mode="eager"), "product_pipeline_invalid"),
        (dict(durable_batch_size=0), "product_pipeline_invalid"),
        (dict(reader_worker_count=True), "product_pipeline_invalid"),
    ]
    for change, code in bad:
        with pytest.raises(InputError) as info:
            generate_psr(dem_path, horizons, tmp_path / "x.tif", sun_vectors_m=sun, backend="cpu",
                         **change)
        assert info.value.code == code, change
    # a transform that changes the dtype it promised fails the calculation
    with pytest.raises(ProductGenerationError) as info:
        generate_psr(dem_path, horizons, tmp_path / "y.tif", sun_vectors_m=sun, backend="cpu",
                     output_transform=lambda b: b.astype(np.int16), output_dtype=np.uint8)
    assert info.value.code == "psr_calculation_failed"
    # a staged job with another transform identity is refused on resume
    with pytest.raises(OperationCancelledError):
        generate_psr(dem_path, horizons, tmp_path / "t.tif", sun_vectors_m=sun, backend="cpu",
                     output_transform=lambda b: b, output_dtype=np.uint8,
                     output_transform_id="identity", cancellation_requested=lambda: True)
    with pytest.raises(IncompatibleProductJobError):
        generate_psr(dem_path, horizons, tmp_path / "t.tif", sun_vectors_m=sun, backend="cpu",
                     output_transform=lambda b: b, output_dtype=np.uint8,
                     output_transform_id="other")


def test_an_unreadable_tile_is_an_invalid_patch(site, tmp_path):
    """psr_pipeline.py:384-389: a tile that fails to read is written invalid."""
    import shutil

    dem, dem_path, horizons, written, sun = site
    store_dir = tmp_path / "horizons"
    shutil.copytree(horizons, store_dir)
    tile = next(store_dir.rglob("horizon_00000_00128_000.cbin"))
    tile.write_bytes(tile.read_bytes()[:100])  # truncated: the codec refuses it
    out = generate_psr(dem_path, store_dir, tmp_path / "u.tif", sun_vectors_m=sun, backend="cpu",
                       invalid_value=5)
    mask, _ = geotiff.read_geotiff(out)
    assert np.all(mask[0:128, 128:256] == 5)
    assert set(np.unique(mask[0:128, 0:128])) <= {0, 255}


@pytest.mark.parametrize("stage", ["read", "calculate", "write", "valid"])
def test_a_failing_callback_in_any_stage_ends_the_bounded_run(site, tmp_path, stage):
    """The first error of any pipeline thread stops the others and is raised as
    is; the staged product stays for a resume."""
    dem, dem_path, horizons, written, sun = site

    class Boom(RuntimeError):
        pass

    def on_event(event):
        if event.stage == stage:
            raise Boom(stage)

    with pytest.raises(Boom):
        generate_psr(dem_path, horizons, tmp_path / f"{stage}.tif", sun_vectors_m=sun,
                     backend="cpu", progress_event_callback=on_event, reader_worker_count=2,
                     decoded_horizon_capacity=2)
    assert not (tmp_path / f"{stage}.tif").exists()
    assert len(list(tmp_path.glob(f".{stage}.tif.lunarscout-partial.*"))) == 3
