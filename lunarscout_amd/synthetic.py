"""Seeded synthetic inputs of the horizon and lighting workloads.

* ``fbm_crater_dem``: float32 fBm relief (spectral synthesis, Hurst exponent H,
  RMS-normalised) plus paraboloid craters with log-uniform radii and a fixed
  depth/diameter ratio, shifted to zero mean, on a south-polar stereographic
  grid centred on a site.
* ``sun_vectors`` / ``earth_vectors``: Moon-ME body positions at a fixed
  cadence; the Sun's elevation at the site follows a seeded two-tone signal
  inside [el_min, el_max], its azimuth turns once per synodic month; the Earth
  librates in longitude and latitude around the sub-Earth point.

The workload helpers (``site_dems`` ...) name the standard sizes.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .geometry import DemGrid, ProjectionParameters, local_direction_vector, project_stereographic

MOON_RADIUS_M = 1_737_400.0
SUN_DISTANCE_M = 1.496e11
EARTH_DISTANCE_M = 3.844e8
SYNODIC_MONTH_H = 29.530589 * 24.0


def site_xy(latitude_deg: float = -89.5, longitude_deg: float = 0.0) -> tuple[float, float]:
    """Projected (x, y) of a site in the south-polar stereographic CRS."""
    return project_stereographic(math.radians(latitude_deg), math.radians(longitude_deg),
                                 ProjectionParameters.south_polar())


def _fbm(size: int, resolution_m: float, hurst: float, rms_m: float,
         rng: np.random.Generator) -> np.ndarray:
    ky = np.fft.fftfreq(size, d=resolution_m)
    kx = np.fft.rfftfreq(size, d=resolution_m)
    k = np.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2)
    k[0, 0] = 1.0
    amplitude = k ** (-(hurst + 1.0))
    amplitude[0, 0] = 0.0
    spectrum = amplitude * (rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape))
    field = np.fft.irfft2(spectrum, s=(size, size))
    field -= field.mean()
    field *= rms_m / max(float(field.std()), 1e-30)
    return field


def _add_craters(z: np.ndarray, resolution_m: float, count: int, r_min: float, r_max: float,
                 depth_ratio: float, rng: np.random.Generator) -> None:
    size = z.shape[0]
    radii = np.exp(rng.uniform(math.log(r_min), math.log(r_max), count))
    centres = rng.uniform(0, size * resolution_m, (count, 2))
    for r, (cx, cy) in zip(radii, centres):
        depth = depth_ratio * 2.0 * r
        c0 = max(int((cx - r) / resolution_m), 0)
        c1 = min(int((cx + r) / resolution_m) + 2, size)
        r0 = max(int((cy - r) / resolution_m), 0)
        r1 = min(int((cy + r) / resolution_m) + 2, size)
        if c0 >= c1 or r0 >= r1:
            continue
        xs = (np.arange(c0, c1) + 0.5) * resolution_m - cx
        ys = (np.arange(r0, r1) + 0.5) * resolution_m - cy
        rho2 = (ys[:, None] ** 2 + xs[None, :] ** 2) / (r * r)
        bowl = np.where(rho2 < 1.0, depth * (rho2 - 1.0), 0.0)
        z[r0:r1, c0:c1] += bowl


def fbm_crater_dem(size: int, resolution_m: float, seed: int, *, hurst: float = 0.8,
                   rms_m: float = 300.0, craters: int = 40, crater_radius_m=(50.0, 3000.0),
                   depth_ratio: float = 0.2, centre_xy: tuple[float, float] | None = None
                   ) -> DemGrid:
    """A size x size float32 DEM at ``resolution_m`` centred on ``centre_xy``."""
    rng = np.random.default_rng(seed)
    z = _fbm(size, resolution_m, hurst, rms_m, rng)
    _add_craters(z, resolution_m, craters, crater_radius_m[0], crater_radius_m[1], depth_ratio, rng)
    z -= z.mean()  # every DEM of a site is centred on the datum
    cx, cy = centre_xy if centre_xy is not None else site_xy()
    half = size * resolution_m / 2.0
    transform = np.array((cx - half, resolution_m, 0.0, cy + half, 0.0, -resolution_m),
                         dtype=np.float64)
    return DemGrid(np.ascontiguousarray(z, dtype=np.float32), transform,
                   ProjectionParameters.south_polar(MOON_RADIUS_M))


def sun_vectors(dem: DemGrid, count: int, step_hours: float, seed: int, *,
                el_min: float = -2.0, el_max: float = 3.0,
                row: int | None = None, column: int | None = None) -> np.ndarray:
    """Moon-ME Sun positions [count, 3] (m) seen from a DEM pixel (default: centre)."""
    rng = np.random.default_rng(seed)
    row = dem.height // 2 if row is None else row
    column = dem.width // 2 if column is None else column
    t = np.arange(count, dtype=np.float64) * step_hours
    phase = rng.uniform(0, 2 * np.pi, 3)
    mid, amp = (el_max + el_min) / 2.0, (el_max - el_min) / 2.0
    el = mid + amp * (0.6 * np.sin(2 * np.pi * t / (27.321661 * 24) + phase[0]) +
                      0.4 * np.sin(2 * np.pi * t / (346.62 * 24) + phase[1]))
    az = np.degrees(phase[2]) + 360.0 * t / SYNODIC_MONTH_H
    return np.stack([local_direction_vector(dem, row, column, a % 360.0, e, SUN_DISTANCE_M)
                     for a, e in zip(az, el)])


def earth_vectors(count: int, step_hours: float, seed: int) -> np.ndarray:
    """Moon-ME Earth positions [count, 3] (m): libration of +-7.9 deg lon, +-6.7 deg lat."""
    rng = np.random.default_rng(seed)
    t = np.arange(count, dtype=np.float64) * step_hours
    p = rng.uniform(0, 2 * np.pi, 2)
    lon = np.radians(7.9) * np.sin(2 * np.pi * t / (27.554550 * 24) + p[0])
    lat = np.radians(6.7) * np.sin(2 * np.pi * t / (27.212221 * 24) + p[1])
    return EARTH_DISTANCE_M * np.stack(
        (np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)), axis=1)


@dataclass(frozen=True)
class SiteConfig:
    """Named sizes of the standard workloads."""

    primary_size: int
    primary_resolution_m: float
    primary_seed: int
    regional_size: int | None = None
    regional_resolution_m: float = 60.0
    regional_seed: int | None = None


ONE_TILE = SiteConfig(1024, 20.0, 0)                  # 1 DEM, central tile (512, 512)
FULL_SITE = SiteConfig(2048, 5.0, 1, 4096, 60.0, 2)   # 256 tiles, 2 DEM passes
ADJOINT_SITE = SiteConfig(512, 10.0, 7, 1024, 60.0, 8)


def site_dems(config: SiteConfig, scale: int = 1) -> list[DemGrid]:
    """Primary (+ regional) DEMs of a workload; ``scale`` > 1 shrinks the grids
    by that factor at the same resolution (small parity cases)."""
    dems = [fbm_crater_dem(config.primary_size // scale, config.primary_resolution_m,
                           config.primary_seed)]
    if config.regional_size:
        dems.append(fbm_crater_dem(config.regional_size // scale, config.regional_resolution_m,
                                   config.regional_seed))
    return dems


@dataclass(frozen=True)
class LandingSite:
    """One landing-site patch of the batched workload."""

    index: int
    seed: int
    tile_count: int     # drawn log-uniform in [1, 64]
    tiles_x: int        # ceil(sqrt(n)) tile columns
    tiles_y: int        # ceil(n / tiles_x) tile rows
    resolution_m: float
    centre_xy: tuple[float, float]


def landing_sites(count: int = 48, first_seed: int = 100, max_tiles: int = 64,
                  offset_km: float = 60.0) -> list[LandingSite]:
    """The uneven-site workload: tile counts n = round(exp(U[0, ln max_tiles])),
    patch DEM ceil(sqrt(n)) x ceil(n / ceil(sqrt(n))) tiles, resolution U[5, 20] m,
    seeds first_seed + i, centres within +-offset_km of the 89.5 S site."""
    out = []
    cx, cy = site_xy()
    for i in range(count):
        rng = np.random.default_rng(first_seed + i)
        n = int(round(math.exp(rng.uniform(0.0, math.log(max_tiles)))))
        tx = int(math.ceil(math.sqrt(n)))
        ty = int(math.ceil(n / tx))
        res = float(rng.uniform(5.0, 20.0))
        dx, dy = rng.uniform(-offset_km * 1000.0, offset_km * 1000.0, 2)
        out.append(LandingSite(i, first_seed + i, n, tx, ty, res, (cx + dx, cy + dy)))
    return out


def landing_site_dem(site: LandingSite) -> DemGrid:
    """The site's patch DEM: (tiles_y * 128) x (tiles_x * 128) px centred on the site."""
    size = 128 * max(site.tiles_x, site.tiles_y)
    dem = fbm_crater_dem(size, site.resolution_m, site.seed, centre_xy=site.centre_xy)
    h, w = 128 * site.tiles_y, 128 * site.tiles_x
    half_w, half_h = w * site.resolution_m / 2.0, h * site.resolution_m / 2.0
    transform = np.array((site.centre_xy[0] - half_w, site.resolution_m, 0.0,
                          site.centre_xy[1] + half_h, 0.0, -site.resolution_m), dtype=np.float64)
    return DemGrid(np.ascontiguousarray(dem.elevation_m[:h, :w]), transform, dem.projection)
