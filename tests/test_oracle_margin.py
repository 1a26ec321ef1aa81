"""Closed-form pins of the oracle's pixel frame, body margin and visibility fraction.

The margin is el - lerp(horizon, az * N / 360) with the bin truncated and the
upper bin wrapped past N - 1 (lightmap.py:116-121, lightmap_cpu.py:121-143,
psr.py:127-134); the Earth-visibility fraction is the mean over times of
[margin >= 0]. The reference checks its margin signals only CPU against CUDA
(tests/numba_horizon/test_phase6b_lightmap.py:215-283), so these tests fix the
oracle (oracle/lsh_oracle.cpp lsho_lighting_patch, horizon_at, azimuth_elevation)
against values the geometry determines:

* body vectors are built from the pixel's geometric east / north / up
  directions (derivatives of the position on the sphere, not the frame the
  oracle forms), so the recovered azimuth and elevation pin the frame's axes,
  the clockwise-from-north sign and the wrap into [0, 360];
* a constant horizon gives margin = el - c exactly;
* a horizon linear in the bin index is interpolated exactly, so the margin is
  el - (a + k * az * N / 360) away from the wrap, and in the last bin the
  interpolation runs from bin N - 1 to bin 0;
* a constructed series of elevations around a constant horizon gives a known
  visible count.

A sign or axis error in the frame, a dropped wrap, a rounding instead of a
truncation of the bin or a miscounted visibility would fail one of them. A
margin of exactly 0 (where the inclusive >= differs from >) cannot be built
from float64 geometry, so the inclusive comparison itself stays pinned only by
the reference's mission-duration rule it follows. The same checks run on the
host build of the product kernel (the "cpu" backend).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from lunarscout_amd.geometry import (DemGrid, ProjectionParameters, inverse_stereographic,
                                     project_stereographic)
from lunarscout_amd.lighting import CpuLightingSession

N = 1440
SUN_DISTANCE_M = 1.496e11
EARTH_DISTANCE_M = 3.844e8


def _site_dem(height_m: float = 120.0) -> DemGrid:
    # 16 x 16 px at 20 m around (0, 15162 m): the 89.5 S site of the polar configurations
    transform = np.array([-160.0, 20.0, 0.0, 15162.0 + 160.0, 0.0, -20.0], dtype=np.float64)
    return DemGrid(np.full((16, 16), height_m, dtype=np.float32), transform,
                   ProjectionParameters.south_polar())


def _geometric_basis(dem: DemGrid, row: int, column: int):
    """Position and east / north / up unit vectors from derivatives of the sphere point."""
    x, y = dem.pixel_to_crs(column, row)
    lat, lon = inverse_stereographic(x, y, dem.projection)
    back = project_stereographic(lat, lon, dem.projection)
    assert abs(back[0] - x) < 1e-6 and abs(back[1] - y) < 1e-6
    radius = dem.projection.radius_m + float(dem.elevation_m[row, column])

    def point(la, lo):
        return radius * np.array([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)])

    h = 1e-6
    north = point(lat + h, lon) - point(lat - h, lon)
    east = point(lat, lon + h) - point(lat, lon - h)
    p = point(lat, lon)
    return p, east / np.linalg.norm(east), north / np.linalg.norm(north), p / np.linalg.norm(p)


def _vectors(dem, row, column, az_deg, el_deg, distance):
    p, east, north, up = _geometric_basis(dem, row, column)
    az, el = np.deg2rad(np.asarray(az_deg, float)), np.deg2rad(np.asarray(el_deg, float))
    d = (np.cos(el) * np.cos(az))[:, None] * north + (np.cos(el) * np.sin(az))[:, None] * east
    d = d + np.sin(el)[:, None] * up
    return p + distance * d


def _margins(horizon_row, vectors, dem, *, backend="oracle"):
    tile = np.ascontiguousarray(np.broadcast_to(np.asarray(horizon_row, np.float32), (128, 128, N)))
    if backend == "oracle":
        out = oracle.lighting_patch(dem, tile, vectors, tile_y=0, tile_x=0, valid_height=1,
                                    valid_width=1)
        return out["margin"][:, 0, 0].astype(np.float64), float(out["visible_fraction"][0, 0])
    cpu = CpuLightingSession(threads=1)
    margin = np.stack(list(cpu.iter_patch_margin_tiles(dem, tile, vectors, tile_y=0, tile_x=0,
                                                       valid_height=1, valid_width=1)))
    return margin[:, 0, 0].astype(np.float64), None


AZIMUTHS = np.array([0.0, 0.1, 12.34, 89.999, 90.0, 135.0, 180.0, 225.5, 270.0, 300.125,
                     359.0, 359.74])
ELEVATIONS = np.array([-2.0, 0.0, 1.5, 3.0, 7.25, -0.5, 0.27, 2.0, 45.0, -1.0, 0.01, 5.0])


@pytest.mark.parametrize("row,column", [(0, 0), (7, 11), (15, 15)])
@pytest.mark.parametrize("distance", [SUN_DISTANCE_M, EARTH_DISTANCE_M])
def test_frame_recovers_geometric_azimuth_and_elevation(row, column, distance):
    dem = _site_dem()
    v = _vectors(dem, row, column, AZIMUTHS, ELEVATIONS, distance)
    az, el = oracle.azimuth_elevation(dem, row, column, v)
    np.testing.assert_allclose(el, ELEVATIONS, atol=1e-7)
    err = (az - AZIMUTHS + 180.0) % 360.0 - 180.0
    assert np.abs(err).max() < 1e-7
    # atan2 % 360 (psr.py:56): a direction a rounding west of north gives 360.0 itself
    assert np.all((az >= 0.0) & (az <= 360.0))


@pytest.mark.parametrize("backend", ["oracle", "cpu"])
def test_constant_horizon_margin_is_elevation_minus_horizon(backend):
    dem = _site_dem()
    c = 1.25
    v = _vectors(dem, 0, 0, AZIMUTHS, ELEVATIONS, EARTH_DISTANCE_M)
    margin, _ = _margins(np.full(N, c), v, dem, backend=backend)
    np.testing.assert_allclose(margin, ELEVATIONS - c, atol=2e-6)


@pytest.mark.parametrize("backend", ["oracle", "cpu"])
def test_linear_horizon_is_interpolated_exactly_with_truncated_bins(backend):
    dem = _site_dem()
    a, k = -3.0, 0.004  # hz[b] = a + k b degrees, stored as float32
    row = (a + k * np.arange(N)).astype(np.float32)
    az = np.array([0.0, 0.1, 0.2499, 0.25, 33.3, 123.456, 200.01, 359.5, 359.7])
    el = np.full(az.size, 2.0)
    v = _vectors(dem, 0, 0, az, el, EARTH_DISTANCE_M)
    margin, _ = _margins(row, v, dem, backend=backend)
    true_az, true_el = oracle.azimuth_elevation(dem, 0, 0, v)
    pos = true_az * (N / 360.0)
    lower = np.floor(pos).astype(int) % N  # due north may come back as 360.0: bin 0
    f = pos - np.floor(pos)
    expected = true_el - (row[lower].astype(np.float64)
                          + f * (row[(lower + 1) % N].astype(np.float64) - row[lower]))
    np.testing.assert_allclose(margin, expected, atol=2e-6)
    # and against the closed form of the line itself (float32 storage of a + k b)
    # outside the last bin, where the interpolation wraps to bin 0 instead
    inside = pos < N - 1
    assert inside.sum() >= 7
    np.testing.assert_allclose(margin[inside], 2.0 - (a + k * pos[inside]), atol=1e-5)


@pytest.mark.parametrize("backend", ["oracle", "cpu"])
def test_last_bin_interpolates_towards_bin_zero(backend):
    dem = _site_dem()
    row = np.zeros(N, dtype=np.float32)
    row[N - 1], row[0] = 4.0, 8.0
    az = np.array([359.75, 359.8, 359.9, 359.99])
    v = _vectors(dem, 0, 0, az, np.zeros(az.size), EARTH_DISTANCE_M)
    margin, _ = _margins(row, v, dem, backend=backend)
    true_az, true_el = oracle.azimuth_elevation(dem, 0, 0, v)
    f = true_az * 4.0 - (N - 1)
    assert np.all((f >= 0.0) & (f < 1.0))
    # a dropped wrap (upper bin N) or a rounded bin would give another horizon here
    np.testing.assert_allclose(margin, true_el - (4.0 + 4.0 * f), atol=2e-6)


def test_visibility_fraction_counts_times_with_margin_at_least_zero():
    dem = _site_dem()
    el = np.array([-3.0, 0.5, 1.2, 5.0, 0.99, 1.01, 2.0, -1.0])
    az = np.linspace(10.0, 350.0, el.size)
    v = _vectors(dem, 3, 4, az, el, EARTH_DISTANCE_M)
    # tile origin (0, 0): pixel (row 3, column 4) of a 4 x 5 valid patch
    tile = np.ascontiguousarray(np.broadcast_to(np.float32(1.0), (128, 128, N)))
    out = oracle.lighting_patch(dem, tile, v, tile_y=0, tile_x=0, valid_height=4,
                                valid_width=5)
    margin = out["margin"][:, 3, 4]
    np.testing.assert_allclose(margin, el - 1.0, atol=2e-6)
    assert out["visible_fraction"][3, 4] == pytest.approx(4.0 / 8.0, abs=0)
