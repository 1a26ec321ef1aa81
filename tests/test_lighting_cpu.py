"""The "cpu" lighting backend (host build of the lighting kernel) and backend selection.

Bars as for the GPU path: C# BuilderSunFraction bytes and C# PSR tiles
byte-exact; sun fraction within 1e-3 and Earth margin within 1e-4 of the
float64 oracle; PSR identical except near the threshold.
"""

from __future__ import annotations

import base64

import numpy as np
import pytest

import oracle
from tests import fixtures as fx
from lunarscout_amd import synthetic
from lunarscout_amd.geometry import DemGrid, ProjectionParameters, local_direction_vector
from lunarscout_amd.lighting import CpuLightingSession, select_lighting_session


def _unit_dem():
    return DemGrid(np.zeros((1, 1), np.float32),
                   np.asarray((1000.0, 1.0, 0.0, -1000.0, 0.0, -1.0), np.float64),
                   ProjectionParameters.south_polar())


def test_sun_fraction_bytes_match_csharp_builder():
    dem, cpu = _unit_dem(), CpuLightingSession(threads=2)
    for case in fx.lightmap_csharp()["cases"]:
        horizon = fx.decode_f32(case["horizons_float32_base64"])
        tile = np.broadcast_to(horizon, (128, 128, horizon.size))
        vectors = np.stack([local_direction_vector(dem, 0, 0, s["azimuth_deg"], s["elevation_deg"],
                                                   1.5e11) for s in case["samples"]])
        got = np.stack(list(cpu.iter_patch_tiles(dem, tile, vectors, tile_y=0, tile_x=0,
                                                 valid_height=1, valid_width=1)))[:, 0, 0]
        np.testing.assert_array_equal(got, [s["encoded_byte"] for s in case["samples"]],
                                      err_msg=case["name"])


def test_psr_matches_csharp_kernel_fixture():
    art = fx.psr_csharp()
    p = art["projection"]
    x = np.arange(128, dtype=np.float64)[None, :]
    y = np.arange(128, dtype=np.float64)[:, None]
    dem = DemGrid(np.ascontiguousarray((x - y) * 0.1, dtype=np.float32),
                  np.asarray(art["geotransform"], np.float64),
                  ProjectionParameters(p["radius_m"], p["latitude_origin_rad"],
                                       p["longitude_origin_rad"], p["scale"],
                                       p["false_easting_m"], p["false_northing_m"]))
    pixel = np.arange(128 * 128, dtype=np.int32)[:, None]
    azimuth = np.arange(1440, dtype=np.int32)[None, :]
    cpu = CpuLightingSession()
    for case in art["cases"]:
        if case["name"] == "constant_shadow":
            hz = np.full((128 * 128, 1440), 0.3, np.float32)
        else:
            hz = (np.float32(0.45) + np.float32(0.01) * (pixel % 23)
                  + np.float32(0.0001) * (azimuth % 17)).astype(np.float32)
            if case["name"] == "compressed_quantized_mixed":
                scaled = hz * np.float32(32767.0 / 50.0)
                hz = (np.floor(scaled + np.float32(0.5)).astype(np.int16)
                      * np.float32(50.0 / 32767.0)).astype(np.float32)
        expected = np.frombuffer(base64.b64decode(case["output_base64"]), dtype=np.uint8).reshape(128, 128)
        got = cpu.compute_patch(dem, hz.reshape(128, 128, 1440),
                                np.asarray(case["sun_vectors_m"], np.float64).reshape(-1, 3),
                                tile_y=0, tile_x=0)
        np.testing.assert_array_equal(got, expected, err_msg=case["name"])


def test_streamed_lighting_matches_the_float64_oracle():
    dem = synthetic.fbm_crater_dem(256, 5.0, 1)
    rng = np.random.default_rng(0)
    az = np.radians(np.arange(1440) * 0.25)
    hz = (rng.uniform(-1.5, 5.5, (128, 128, 1)) + 1.5 * np.sin(3 * az)[None, None, :]).astype(np.float32)
    sun = synthetic.sun_vectors(dem, 48, 6.0, seed=3)
    earth = synthetic.earth_vectors(48, 6.0, seed=4)
    reduced, _ = oracle.psr_reduce(dem, sun)
    out = CpuLightingSession(threads=4).lighting_tiles(
        dem, [(0, 128)], hz.reshape(1, 16384, 1440), sun=sun, earth=earth, psr_vectors=reduced,
        sun_fraction=True, earth_margin=True, earth_visible=True, psr=True)
    ref = oracle.lighting_patch(dem, hz, sun, tile_y=128, tile_x=0)
    assert np.abs(out["sun_fraction"][0].reshape(-1, 128, 128) - ref["fraction"]).max() <= 1e-3
    eref = oracle.lighting_patch(dem, hz, earth, tile_y=128, tile_x=0)
    assert np.abs(out["earth_margin"][0].reshape(-1, 128, 128) - eref["margin"]).max() <= 1e-4
    mask, max_margin = oracle.psr_patch(dem, hz, reduced, tile_y=128, tile_x=0)
    differ = out["psr"][0].reshape(128, 128) != mask
    near = np.abs(max_margin - (-float(np.float32(0.545)) / 2)) <= 1e-3
    assert not np.any(differ & ~near)
    assert 0 < int((mask == 255).sum()) < 128 * 128  # the case exercises both classes


def test_backend_selection():
    session, selected = select_lighting_session("cpu")
    assert selected == "cpu" and isinstance(session, CpuLightingSession)
    with pytest.raises(ValueError):
        select_lighting_session("cuda")
    from lunarscout_amd import _lib

    if _lib.device_count() == 0:
        _, selected = select_lighting_session("auto")
        assert selected == "cpu"
        from lunarscout_amd.errors import HipBackendError

        with pytest.raises(HipBackendError):
            select_lighting_session("hip")


def test_disk_walk_equals_the_reference_walk_bit_for_bit_on_steep_rows():
    """The product's slice walk (with its whole-disk lit / dark exits) against the
    oracle's restatement of _sun_fraction_reference (lightmap.py:32-73) in float32.

    Rows are random walks with up to +-4 degrees between bins, each pixel its own
    row; the Sun sits at float32-exact azimuths (including the truncated negative
    positions below 0.395 deg and the wrap past 359.6 deg) and at elevations
    within a disk of the local horizon, fully above and fully below it. The
    fractions must be identical.
    """
    rng = np.random.default_rng(21)
    dem = DemGrid(np.zeros((128, 128), np.float32),
                  np.asarray((0.0, 1.0, 0.0, 15162.0 + 64.0, 0.0, -1.0), np.float64),
                  ProjectionParameters.south_polar())
    steps = rng.uniform(-4.0, 4.0, (128, 128, 1440)).astype(np.float32)
    rows = np.cumsum(steps, axis=-1, dtype=np.float32)
    rows = (rows - rows.mean(axis=-1, keepdims=True)).astype(np.float32)
    az = np.concatenate([rng.uniform(0.0, 0.4, 40), rng.uniform(359.5, 360.0, 40),
                         rng.uniform(0.0, 360.0, 120)]).astype(np.float32)
    pixel = (rng.integers(0, 128), rng.integers(0, 128))
    row = rows[pixel]
    bins = (az.astype(np.float64) * 4).astype(int) % 1440
    el = (row[bins].astype(np.float64) + rng.choice([-6.0, -1.0, -0.2, 0.0, 0.1, 0.3, 1.0, 6.0], az.size)
          + rng.uniform(-0.1, 0.1, az.size)).astype(np.float32)
    vectors = np.stack([local_direction_vector(dem, pixel[0], pixel[1], float(a), float(e), 1.5e11)
                        for a, e in zip(az, el)])
    got_az, got_el = oracle.azimuth_elevation(dem, pixel[0], pixel[1], vectors)
    assert np.array_equal(got_az.astype(np.float32), az) and np.array_equal(got_el.astype(np.float32), el)
    cpu = CpuLightingSession(threads=2)
    fractions = np.stack(list(cpu.iter_patch_fraction_tiles(dem, rows, vectors, tile_y=0, tile_x=0)))
    got = fractions[:, pixel[0], pixel[1]]
    want = np.array([oracle.sun_fraction(row, float(a), float(e), float32=True)
                     for a, e in zip(az, el)], dtype=np.float32)
    assert np.count_nonzero(want == 1.0) > 10 and np.count_nonzero(want == 0.0) > 10
    np.testing.assert_array_equal(got, want)


def _fake_device_buffer(shape, dtype):
    """A DeviceBuffer descriptor with no device memory (the checks read only shape/dtype)."""
    from lunarscout_amd._lib import DeviceBuffer

    buf = DeviceBuffer.__new__(DeviceBuffer)
    buf.session, buf.ptr = None, 0
    buf.dtype, buf.shape = np.dtype(dtype), tuple(shape)
    buf.nbytes = int(np.prod(shape)) * buf.dtype.itemsize
    buf.free = lambda: None
    return buf


def test_device_inputs_are_checked_before_any_kernel_reads_them():
    from lunarscout_amd.errors import InputError
    from lunarscout_amd.lighting import _check_device_horizons, _check_device_vectors

    assert _check_device_horizons(_fake_device_buffer((3, 16384, 1440), np.float32), 3, False) == 1440
    assert _check_device_horizons(_fake_device_buffer((2, 720, 16384), np.float32), 2, True) == 720
    with pytest.raises(InputError, match="fewer tiles"):
        _check_device_horizons(_fake_device_buffer((2, 16384, 1440), np.float32), 3, False)
    with pytest.raises(InputError, match="float32"):
        _check_device_horizons(_fake_device_buffer((3, 16384, 1440), np.float64), 3, False)
    with pytest.raises(InputError, match="16384 pixels"):
        _check_device_horizons(_fake_device_buffer((3, 1440, 16384), np.float32), 3, False)
    assert _check_device_vectors(_fake_device_buffer((40, 3), np.float64), "sun", 40) == 40
    with pytest.raises(InputError, match="fewer entries"):
        _check_device_vectors(_fake_device_buffer((39, 3), np.float64), "sun", 40)
    with pytest.raises(InputError, match="float64"):
        _check_device_vectors(_fake_device_buffer((40, 3), np.float32), "earth")


def test_host_lighting_reads_either_layout_to_the_same_bits():
    """The host build walks [pixel][azimuth] and [azimuth][pixel] rows alike."""
    dem = synthetic.fbm_crater_dem(256, 20.0, 12)
    rng = np.random.default_rng(12)
    az = np.arange(1440) * (2 * np.pi / 1440)
    # smooth rows with steep walls, so the disk walk and the lit/dark exits both run
    base = rng.normal(0.0, 2.0, (2, 16384, 1)) + 3.0 * np.sin(az[None, None, :] * rng.integers(
        1, 5, (2, 16384, 1)) + rng.uniform(0, 6.3, (2, 16384, 1)))
    horizons = np.maximum(base, -50.0).astype(np.float32)
    tiles = [(0, 0), (128, 128)]
    sun = synthetic.sun_vectors(dem, 48, 2.0, seed=13)
    earth = synthetic.earth_vectors(48, 2.0, seed=14)
    cpu = CpuLightingSession(threads=4)
    names = ("sun_bytes", "sun_fraction", "sun_mean", "earth_margin", "earth_visible", "psr",
             "psr_margin")
    got = [cpu.lighting_tiles(dem, tiles, data, sun=sun, earth=earth, psr_vectors=sun[:24],
                              azimuth_major=major, **{k: True for k in names})
           for major, data in ((False, horizons),
                               (True, np.ascontiguousarray(horizons.transpose(0, 2, 1))))]
    for k in names:
        np.testing.assert_array_equal(got[0][k], got[1][k], err_msg=k)
    assert 0.0 < float(got[0]["sun_fraction"].mean()) < 1.0


@pytest.mark.parametrize("bins", [720, 1440, 2880])
def test_disk_walk_stays_in_its_row_for_any_bin_count(bins):
    """The slice walk of an N-bin row reads only that row, at every azimuth.

    Every pixel's row is a constant -10 degree horizon and the rows either side
    of the sampled pixels hold +80 degrees, so a read past the row's end (the
    wrap past 359.x deg) or before its start would darken a Sun at 0 degrees
    elevation. With the rows in bounds the disk is fully lit at every azimuth,
    for N = 720, 1440 and 2880 alike.
    """
    dem = DemGrid(np.zeros((128, 128), np.float32),
                  np.asarray((0.0, 1.0, 0.0, 15162.0 + 64.0, 0.0, -1.0), np.float64),
                  ProjectionParameters.south_polar())
    rows = np.full((128, 128, bins), 80.0, np.float32)
    pixels = [(5, 7), (64, 64), (127, 126)]
    for y, x in pixels:
        rows[y, x] = -10.0
    az = np.concatenate([np.linspace(0.0, 0.5, 6), np.linspace(359.4, 359.99, 8),
                         np.linspace(10.0, 350.0, 9)]).astype(np.float32)
    cpu = CpuLightingSession(threads=2)
    for y, x in pixels:
        vectors = np.stack([local_direction_vector(dem, y, x, float(a), 0.0, 1.5e11) for a in az])
        fractions = np.stack(list(cpu.iter_patch_fraction_tiles(dem, rows, vectors, tile_y=0,
                                                                tile_x=0)))
        np.testing.assert_array_equal(fractions[:, y, x], np.ones(az.size, np.float32))


def test_row_decisions_leave_fractions_visibility_and_mask_unchanged():
    """The visible fraction and the PSR mask asked for alone (whole-row decisions
    on) equal those derived with the Earth margins and the PSR margin asked for
    too (decisions off), and the sun fractions equal the oracle's float32 walk."""
    dem = synthetic.fbm_crater_dem(256, 5.0, 1)
    rng = np.random.default_rng(2)
    az = np.radians(np.arange(1440) * 0.25)
    hz = (rng.uniform(-2.5, 4.0, (128, 128, 1)) + 1.2 * np.sin(3 * az)[None, None, :]).astype(np.float32)
    hz[:24] += np.float32(8.0)  # pixels that no Sun vector reaches (PSR)
    # series of at least 512 times (shorter ones take no decisions)
    sun = synthetic.sun_vectors(dem, 600, 4.0, seed=8)
    earth = synthetic.earth_vectors(600, 4.0, seed=9)
    reduced, _ = oracle.psr_reduce(dem, sun)
    rows = hz.reshape(1, 16384, 1440)
    common = dict(sun=sun, earth=earth, psr_vectors=reduced)
    alone = CpuLightingSession(threads=4).lighting_tiles(
        dem, [(0, 128)], rows, sun_fraction=True, earth_visible=True, psr=True, **common)
    full = CpuLightingSession(threads=4).lighting_tiles(
        dem, [(0, 128)], rows, sun_fraction=True, earth_margin=True, earth_visible=True, psr=True,
        psr_margin=True, **common)
    np.testing.assert_array_equal(alone["earth_visible"], full["earth_visible"])
    np.testing.assert_array_equal(alone["psr"], full["psr"])
    ref = oracle.lighting_patch(dem, hz, sun, tile_y=128, tile_x=0, float32_slices=True)
    fraction = alone["sun_fraction"][0].reshape(-1, 128, 128)
    np.testing.assert_array_equal(fraction, ref["fraction"].astype(np.float32))
    decided = np.mean((fraction == 0) | (fraction == 1))
    assert 0.2 < decided < 1.0, decided  # the case exercises both decided and walked times
    assert 0 < int((alone["psr"] == 255).sum()) < 16384


def test_row_decisions_see_the_walk_extrapolate_left_of_bin_zero():
    """A disk centred within half a bin plus the disk radius of azimuth 0 starts
    its walk left of bin 0 and extrapolates from bins (0, 1): a row whose lowest
    bin is 1 deg can still show the disk an extrapolated 0 deg horizon, and one
    whose highest is 5 deg can hide it behind an extrapolated 6 deg. The
    fractions (decided or walked) equal the oracle's float32 walk."""
    from lunarscout_amd.geometry import local_direction_vector

    dem = synthetic.fbm_crater_dem(256, 5.0, 1)
    row, col = 128 + 40, 40
    cases = [((1.0, 2.0, 1.0), 0.3), ((4.0, 2.0, 5.0), 5.5)]  # (h[0], h[1], h[1439]), elevation
    for (h0, h1, hn), el in cases:
        rowv = np.full(1440, 5.0 if el < 3 else 3.0, np.float32)
        rowv[0], rowv[1], rowv[1439] = h0, h1, hn
        hz = np.broadcast_to(rowv, (128, 128, 1440)).astype(np.float32)
        azs = [0.02, 0.1, 0.2, 0.3, 0.38, 0.5, 359.9]
        vec = np.stack([local_direction_vector(dem, row, col, a, el, 1.496e11) for a in azs])
        vec = np.tile(vec, (80, 1))  # 560 times: long enough to take decisions
        out = CpuLightingSession(threads=2).lighting_tiles(
            dem, [(0, 128)], hz.reshape(1, 16384, 1440), sun=vec, sun_fraction=True)
        got = out["sun_fraction"][0].reshape(-1, 128, 128)
        ref = oracle.lighting_patch(dem, hz, vec, tile_y=128, tile_x=0, float32_slices=True)
        np.testing.assert_array_equal(got, ref["fraction"].astype(np.float32))
        partial = got[:, 40, 40]
        assert np.any((partial > 0) & (partial < 1)), partial  # the extrapolated cases
