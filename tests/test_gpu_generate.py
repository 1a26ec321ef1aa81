"""End to end on the GPU: generate_horizons over GeoTIFF DEMs, resume, overwrite.

The tiles it writes are compared with the HIP session's own degrees for the same
tile (the march itself is held to the float64 oracle in test_gpu_horizon.py):
.cbin tiles within half a quantisation step (50/32767/2 deg) plus float32
rounding, .bin tiles exactly, padding at -50 deg.
"""

from __future__ import annotations

import numpy as np
import pytest

from lunarscout_amd import geotiff, synthetic, tiles
from lunarscout_amd.geometry import DemGrid
from lunarscout_amd.horizon import generate_horizons, pyramid_cache_path

pytestmark = pytest.mark.gpu

QUANTUM = float(np.float32(50.0) / np.float32(32767.0))


def _dems(tmp_path):
    primary = synthetic.fbm_crater_dem(512, 20.0, 8)
    primary = DemGrid(np.ascontiguousarray(primary.elevation_m[:200, :300]), primary.geo_transform,
                      primary.projection)
    regional = synthetic.fbm_crater_dem(256, 60.0, 9)  # centred on the same site
    p1 = geotiff.write_dem(tmp_path / "primary.tif", primary)
    p2 = geotiff.write_dem(tmp_path / "regional.tif", regional)
    return [geotiff.load_dem(p1), geotiff.load_dem(p2)], [p1, p2]


def test_generate_horizons_writes_resumes_and_overwrites(tmp_path):
    from lunarscout_amd.hip_backend import HipSession

    dems, paths = _dems(tmp_path)
    out = tmp_path / "horizons"
    fractions, events = [], []
    print("generate_horizons: 300x200 primary + 256^2 regional, 6 tiles", flush=True)
    result = generate_horizons(out, paths, observer_height_m=1.5, progress_callback=fractions.append,
                               progress_event_callback=events.append)
    assert result == out.resolve()
    assert fractions == sorted(fractions) and fractions[-1] == 1.0
    assert {e.backend for e in events} == {"hip"} and events[-1].stage == "complete"
    for dem, path in zip(dems, paths):
        cache = np.fromfile(pyramid_cache_path(path), dtype="<f4")
        session = HipSession(0)
        np.testing.assert_array_equal(cache, session.build_max_pyramid(dem).mips)
        session.close()
    store = tiles.HorizonTileStore(out)
    session = HipSession(0)
    for tx, ty, vw, vh in ((0, 0, 128, 128), (256, 0, 44, 128), (128, 128, 128, 72), (256, 128, 44, 72)):
        path = store.find_existing_path(ty, tx, 1.5)
        assert path is not None and path.suffix == ".cbin" and path.name.endswith("_015.cbin")
        got = tiles.read_horizon_tile(path)
        want = session.horizon_tile(dems, tile_column=tx, tile_row=ty, observer_elevation_m=1.5,
                                    degrees=True)
        # the tile is exactly the codec's image of the session's degrees
        normalized = tiles.normalize_tile(want, valid_width=vw, valid_height=vh)
        expected = tiles.decode_cbin(tiles.encode_cbin(normalized)).reshape(128, 128, 1440)
        np.testing.assert_array_equal(got, expected)
        assert np.all(got[vh:] == np.float32(-50.0)) and np.all(got[:, vw:] == np.float32(-50.0))
        # away from the +-16384-step delta clamp the codec is within half a step
        smooth = np.abs(np.diff(np.clip(normalized, -50, 50), axis=1)).max(axis=1) < 20.0
        err = np.abs(expected.reshape(16384, 1440) - np.clip(normalized, -50, 50))[smooth]
        assert err.size == 0 or err.max() <= 0.5 * QUANTUM + 8e-6
    session.close()
    # resume: nothing is regenerated
    mtimes = {p: p.stat().st_mtime_ns for p in out.rglob("*.cbin")}
    events.clear()
    generate_horizons(out, paths, observer_height_m=1.5, progress_event_callback=events.append)
    assert {p: p.stat().st_mtime_ns for p in out.rglob("*.cbin")} == mtimes
    assert events[-1].stage == "complete" and events[-1].total == 0
    # overwrite as raw .bin; the bin tile holds the float32 degrees exactly
    generate_horizons(out, paths, observer_height_m=1.5, compress=False, overwrite=True)
    raw = tiles.read_horizon_tile(out / "00000" / "horizon_00000_00000_015.bin")
    session = HipSession(0)
    want = session.horizon_tile(dems, tile_column=0, tile_row=0, observer_elevation_m=1.5,
                                degrees=True).reshape(128, 128, 1440)
    session.close()
    np.testing.assert_array_equal(raw, want)
